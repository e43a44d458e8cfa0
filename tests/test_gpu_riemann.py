"""Pointwise parity of the device exact two-material Riemann solver (m2c_riemann_batch) with the
oracle's restatement of ExactRiemannSolverBase::ComputeRiemannSolution, on seeded random
problems across stiffened-gas pairs (water/air, air/helium-like) in all three directions."""
import numpy as np
import pytest

import m2c_amd as M
from m2c_amd import workloads as W
from oracle import pyoracle as po
from tests import cases

pytestmark = pytest.mark.gpu

MATS = [M.material(M.EOS_SG, [6.12, 3.43e8]), M.material(M.EOS_SG, [1.4, 0.0]), M.material(M.EOS_SG, [1.67, 0.0])]


def problems(n, seed):
    rng = np.random.default_rng(seed)
    idm = rng.integers(0, 3, n).astype(np.int32)
    idp = rng.integers(0, 3, n).astype(np.int32)
    d = rng.integers(0, 3, n).astype(np.int32)
    Vm, Vp = np.zeros((n, 5)), np.zeros((n, 5))
    for V, ids in ((Vm, idm), (Vp, idp)):
        water = ids == 0
        V[:, 0] = np.where(water, 1000.0 * np.exp(rng.uniform(-0.05, 0.2, n)), np.exp(rng.uniform(-2.0, 2.5, n)))
        V[:, 4] = np.where(water, np.exp(rng.uniform(np.log(1e5), np.log(2e9), n)),
                           np.exp(rng.uniform(np.log(1e4), np.log(1e7), n)))
        V[:, 1:4] = rng.uniform(-300.0, 300.0, (n, 3))
    return d, Vm, idm, Vp, idp


def test_riemann_batch_matches_oracle_bitwise():
    settings = dict(nx=4, ny=4, nz=4, box=(0, 1, 0, 1, 0, 1), bc=[M.BC_SYMMETRY] * 6, mats=MATS)
    s = M.Solver(W.config_from(settings))
    o = po.Oracle(cases.oracle_config(settings))
    d, Vm, idm, Vp, idp = problems(3000, 7)
    st, Vs, mid, Vsm, Vsp = s.riemann(d, Vm, idm, Vp, idp)
    bad = []
    nfail = 0
    for t in range(len(d)):
        dd = np.zeros(3)
        dd[d[t]] = 1.0
        ost, oVs, omid, oVsm, oVsp = o.riemann(dd, Vm[t], int(idm[t]), Vp[t], int(idp[t]))
        nfail += ost != 0
        same = (ost == st[t] and omid == mid[t] and np.array_equal(oVs, Vs[t]) and np.array_equal(oVsm, Vsm[t])
                and np.array_equal(oVsp, Vsp[t]))
        if not same:
            bad.append((t, ost, int(st[t]), float(np.max(np.abs(oVsm - Vsm[t]) / (np.abs(oVsm) + 1e-300)))))
    print(f"problems {len(d)}, oracle failures {nfail}, mismatches {len(bad)}: {bad[:8]}", flush=True)
    assert not bad
