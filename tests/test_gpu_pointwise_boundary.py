"""GPU parity of the pointwise drop-in boundary against the oracle, through the C ABI and through the
C++ host layer shaped like the reference's classes:
* ExactRiemannSolverBase::ComputeRiemannSolution along arbitrary unit normals
  (ExactRiemannSolverBase.h:35-39; m2c_riemann_batch takes double *dir);
* VarFcnBase's pointwise set (VarFcnBase.h:62-294; m2c_varfcn_batch) for stiffened gas, NASG,
  Mie-Grueneisen and JWL;
* FluxFcnBase::ComputeNumericalFluxAtMaterialInterface (FluxFcnBase.h:72-76, the LLF form of
  FluxFcnLLF.h:112-185; m2c_interface_flux_batch).
FP64 results must be bit-identical to the oracle's; JWL evaluates exp() and is held to 1e-13 relative
(device and host exp may differ in the last bit)."""
import os
import subprocess

import numpy as np
import pytest

import m2c_amd as M
from m2c_amd import workloads as W
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

WATER_NASG = dict(gam=1.19, pc=6.217e8, b=6.7212e-4, q=-1177788.0)
ALU_MG = dict(rho0=2785.0, c0=5328.0, Gamma0=2.0, s=1.338, e0=0.0)
TNT_JWL = dict(omega=0.28, A1=3.712e11, A2=3.23e9, R1=4.15, R2=0.95, rho0=1630.0)


def materials():
    return [po.sg(1.4, 0.0), po.sg(6.12, 3.43e8), po.eos(po.EOS_NASG, **WATER_NASG),
            po.eos(po.EOS_MG, **ALU_MG), po.eos(po.EOS_JWL, **TNT_JWL)]


def solver_and_oracle():
    mats = [W.material(M.EOS_SG, [1.4, 0.0]), W.material(M.EOS_SG, [6.12, 3.43e8]),
            W.material(M.EOS_NASG, [WATER_NASG[k] for k in ("gam", "pc", "b", "q")]),
            W.material(M.EOS_MG, [ALU_MG[k] for k in ("rho0", "c0", "Gamma0", "s", "e0")]),
            W.material(M.EOS_JWL, [TNT_JWL[k] for k in ("omega", "A1", "A2", "R1", "R2", "rho0")])]
    settings = dict(nx=4, ny=4, nz=4, box=(0, 1, 0, 1, 0, 1), bc=[M.BC_SYMMETRY] * 6, mats=mats)
    s = M.Solver(W.config_from(settings))
    o = po.Oracle(po.make_config(4, 4, 4, (0, 1, 0, 1, 0, 1), [po.BC_SYMMETRY] * 6, materials()))
    return s, o


# states (rho, e, p) inside each material's valid region
STATES = {0: [(1.2, 2.1e5, 1e5), (0.5, 6.0e5, 2e5)], 1: [(1000.0, 4.2e5, 1e5), (1200.0, 5.0e5, 1e9)],
          2: [(1000.0, 3.0e5, 1e6), (1100.0, 6.0e5, 5e8)], 3: [(2785.0, 1.0e5, 1e5), (3000.0, 4.0e5, 2e9)],
          4: [(1630.0, 4.3e6, 2e9), (1000.0, 3.0e6, 8e8)]}


def close(a, b, jwl):
    if not jwl:
        return np.array_equal(a, b)
    return np.allclose(a, b, rtol=1e-13, atol=0)


def test_riemann_along_arbitrary_normals():
    print("case riemann normals", flush=True)
    s, o = solver_and_oracle()
    rng = np.random.default_rng(5)
    n = 400
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    idm = rng.integers(0, 2, n)
    idp = rng.integers(0, 2, n)
    Vm = np.zeros((n, 5))
    Vp = np.zeros((n, 5))
    for V, ids in ((Vm, idm), (Vp, idp)):
        V[:, 0] = np.where(ids == 0, rng.uniform(0.5, 2.0, n), rng.uniform(990.0, 1100.0, n))
        V[:, 1:4] = rng.uniform(-60.0, 60.0, (n, 3))
        V[:, 4] = np.where(ids == 0, rng.uniform(5e4, 3e5, n), rng.uniform(1e5, 1e8, n))
    st, Vs, mid, Vsm, Vsp = s.riemann(d, Vm, idm, Vp, idp)
    s.close()
    nfail = 0
    for t in range(n):
        ost, oVs, omid, oVsm, oVsp = o.riemann(d[t], Vm[t], int(idm[t]), Vp[t], int(idp[t]))
        assert st[t] == ost and mid[t] == omid, t
        assert np.array_equal(Vs[t], oVs) and np.array_equal(Vsm[t], oVsm) and np.array_equal(Vsp[t], oVsp), t
        nfail += ost != 0
    print(f"  {n} problems, {nfail} solver failures, all bit-identical", flush=True)


@pytest.mark.parametrize("mid", range(5))
def test_varfcn_pointwise_set(mid):
    print(f"case varfcn material {mid}", flush=True)
    s, o = solver_and_oracle()
    jwl = mid == 4
    pts = np.array(STATES[mid], dtype=np.float64)
    re = np.ascontiguousarray(pts[:, [0, 1]])
    rp = np.ascontiguousarray(pts[:, [0, 2]])
    out = {op: s.varfcn(op, mid, re)[0] for op in (M.VF_PRESSURE, M.VF_DPDRHO, M.VF_BIG_GAMMA,
                                                   M.VF_SOUND_SPEED_SQUARE, M.VF_SOUND_SPEED)}
    out[M.VF_INTERNAL_ENERGY] = s.varfcn(M.VF_INTERNAL_ENERGY, mid, rp)[0]
    out[M.VF_ENTHALPY] = s.varfcn(M.VF_ENTHALPY, mid, rp)[0]
    rng = np.random.default_rng(mid)
    V = np.zeros((len(pts), 5))
    V[:, 0], V[:, 4] = pts[:, 0], pts[:, 2]
    V[:, 1:4] = rng.uniform(-50, 50, (len(pts), 3))
    U = s.varfcn(M.VF_P2C, mid, V)[0]
    V2 = s.varfcn(M.VF_C2P, mid, U)[0]
    H = s.varfcn(M.VF_TOTAL_ENTHALPY, mid, V)[0]
    mach = s.varfcn(M.VF_MACH, mid, V)[0]
    chk, chkf = s.varfcn(M.VF_CHECK_STATE, mid, V)
    bad = V.copy()
    bad[:, 4] = -1e12
    chkb, _ = s.varfcn(M.VF_CHECK_STATE, mid, bad)
    for t, (rho, e, p) in enumerate(STATES[mid]):
        ref = o.eos_point(rho, e, p, mid)
        for op, k in ((M.VF_PRESSURE, 0), (M.VF_INTERNAL_ENERGY, 1), (M.VF_DPDRHO, 2), (M.VF_BIG_GAMMA, 3),
                      (M.VF_SOUND_SPEED_SQUARE, 4)):
            assert close(out[op][t], ref[k], jwl), (op, out[op][t], ref[k])
        assert close(out[M.VF_SOUND_SPEED][t], np.sqrt(ref[4]), jwl)
        ee = o.eos_point(rho, 0.0, p, mid)[1]
        assert close(out[M.VF_ENTHALPY][t], ee + p / rho, jwl)
        assert close(U[t], o.p2c(V[t], mid), jwl)
        assert close(V2[t], o.c2p(U[t], mid), jwl)
        eV = o.eos_point(V[t, 0], 0.0, V[t, 4], mid)[1]
        Hx = eV + 0.5 * (V[t, 1] * V[t, 1] + V[t, 2] * V[t, 2] + V[t, 3] * V[t, 3]) + V[t, 4] / V[t, 0]
        assert close(H[t], Hx, jwl)
        c = np.sqrt(o.eos_point(V[t, 0], eV, 0.0, mid)[4])
        assert close(mach[t], np.sqrt(V[t, 1] ** 2 + V[t, 2] ** 2 + V[t, 3] ** 2) / c, jwl) or \
            abs(mach[t] - np.sqrt(V[t, 1] * V[t, 1] + V[t, 2] * V[t, 2] + V[t, 3] * V[t, 3]) / c) <= 1e-15 * mach[t]
        assert chk[t] == ref[5] == 0.0 and chkf[t] == 0
        assert chkb[t] == 1.0
    s.close()


def test_varfcn_clip():
    print("case varfcn clip", flush=True)
    mats = [W.material(M.EOS_SG, [1.4, 0.0], rhomin=0.5, pmin=0.2, rhomax=3.0, pmax=10.0)]
    s = M.Solver(W.config_from(dict(nx=4, ny=4, nz=4, box=(0, 1, 0, 1, 0, 1), bc=[M.BC_SYMMETRY] * 6, mats=mats)))
    V = np.array([[0.1, 1.0, 2.0, 3.0, 0.1], [1.0, 0.0, 0.0, 0.0, 1.0], [5.0, 0.0, 0.0, 0.0, 50.0]])
    out, flag = s.varfcn(M.VF_CLIP, 0, V)
    s.close()
    assert np.array_equal(out, [[0.5, 1.0, 2.0, 3.0, 0.2], [1.0, 0.0, 0.0, 0.0, 1.0], [3.0, 0.0, 0.0, 0.0, 10.0]])
    assert list(flag) == [1, 0, 1]


def test_material_interface_flux():
    print("case material interface flux", flush=True)
    s, o = solver_and_oracle()
    rng = np.random.default_rng(9)
    n = 200
    idm = rng.integers(0, 5, n)
    idp = rng.integers(0, 5, n)
    Vm = np.zeros((n, 5))
    Vp = np.zeros((n, 5))
    for V, ids in ((Vm, idm), (Vp, idp)):
        for t in range(n):
            rho, e, p = STATES[int(ids[t])][rng.integers(0, 2)]
            V[t] = [rho * rng.uniform(0.95, 1.05), *rng.uniform(-40, 40, 3), p]
    for d in range(3):
        F = s.interface_flux(d, Vm, idm, Vp, idp)
        for t in range(n):
            ref = o.interface_flux(d, Vm[t], int(idm[t]), Vp[t], int(idp[t]))
            assert close(F[t], ref, idm[t] == 4 or idp[t] == 4), (d, t, F[t], ref)
    s.close()


def test_cpp_host_layer_pointwise(tmp_path):
    # VarFcn, FluxFcn and ExactRiemannSolver of m2c_amd/host/m2c_host.hpp, driven by m2c_demo
    print("case host layer pointwise", flush=True)
    demo = os.path.join(os.path.dirname(M.__file__), "m2c_demo")
    rng = np.random.default_rng(17)
    n = 120
    q = np.zeros((n, 16))
    o = po.Oracle(po.make_config(4, 4, 4, (0, 1, 0, 1, 0, 1), [po.BC_SYMMETRY] * 6, materials()))
    for t in range(n):
        mid = int(rng.integers(0, 5))
        rho, e, p = STATES[mid][int(rng.integers(0, 2))]
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        # the Riemann problem and fluxes pair air / water only (the device solver's materials)
        pair = int(rng.integers(0, 2))
        q[t] = [mid, rho, e, p, *rng.uniform(-30, 30, 3), *d, rng.integers(0, 3),
                1.1 if pair == 0 else 1010.0, 1.5e5 if pair == 0 else 2e6, pair, 0, 0]
    src = tmp_path / "pw.bin"
    q.tofile(str(src))
    out = tmp_path / "out.bin"
    r = subprocess.run([demo, "pointwise", str(src), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = np.fromfile(str(out), dtype=np.float64).reshape(n, 32)
    for t in range(n):
        mid, rho, e, p = int(q[t, 0]), q[t, 1], q[t, 2], q[t, 3]
        jwl = mid == 4
        ref = o.eos_point(rho, e, p, mid)
        assert close(res[t, 0:6], ref[[0, 1, 2, 3, 4, 5]], jwl), t
        V = np.array([rho, q[t, 4], q[t, 5], q[t, 6], p])
        U = o.p2c(V, mid)
        assert close(res[t, 6:11], U, jwl) and close(res[t, 11:16], o.c2p(U, mid), jwl)
        idp = int(q[t, 13])
        Vp = np.array([q[t, 11], -q[t, 4], q[t, 6], q[t, 5], q[t, 12]])
        axis = int(q[t, 10])
        if mid <= 1:
            ost, oVs, omid, oVsm, oVsp = o.riemann(q[t, 7:10], V, mid, Vp, idp)
            assert res[t, 17] == (0 if ost == 0 else 1) and res[t, 18] == omid
            assert np.array_equal(res[t, 19:24], oVsm) and res[t, 29] == oVsp[0] and res[t, 30] == oVsp[4]
            assert res[t, 31] == oVs[0]
        Fref = o.flux(axis, V, Vp, mid) if mid == idp else o.interface_flux(axis, V, mid, Vp, idp)
        assert close(res[t, 24:29], Fref, jwl or idp == 4), t
