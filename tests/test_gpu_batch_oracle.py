"""Batched instances of the 64-instance workload (edges 48..160, log-uniform states, random split
planes, RK2, HLLC + van Albada) against the oracle: a batch of eight full-size instances advanced in one
persistent launch per RK stage with the hipGraph step; in FP64 three of them must be bit-identical to
the oracle (van Albada powers as dq*dq) after three steps, and the mixed form's per-field relative L1
and L-infinity against the oracle are reported and held to the 1e-5 L1 bar."""
import numpy as np
import pytest

import m2c_amd as M
from m2c_amd import workloads as W
from tests import cases

pytestmark = pytest.mark.gpu

SELECT = list(range(8))
CHECK = [0, 3, 6]
STEPS = 3


def rel_err(V, Vo):
    out = {}
    vel1 = np.abs(Vo[..., 1:4]).sum()
    velinf = np.abs(Vo[..., 1:4]).max()
    for f, name in enumerate(("rho", "u", "v", "w", "p")):
        d = np.abs(V[..., f] - Vo[..., f])
        n1 = vel1 if 1 <= f <= 3 else np.abs(Vo[..., f]).sum()
        ninf = velinf if 1 <= f <= 3 else np.abs(Vo[..., f]).max()
        out[name] = (float(d.sum() / n1) if n1 > 0 else 0.0, float(d.max() / ninf) if ninf > 0 else 0.0)
    return out


@pytest.mark.parametrize("prec", [M.PREC_FP64, M.PREC_MIXED])
def test_full_size_batch_against_oracle(prec):
    insts = W.batched_riemann(n=64, select=SELECT)
    print(f"case batch prec {prec}: " + ", ".join(f"{s['nx']}x{s['ny']}x{s['nz']}" for s, _, _ in insts), flush=True)
    b = M.Batch([W.config_from(s, precision=prec) for s, _, _ in insts])
    for i, (s, V0, ID0) in enumerate(insts):
        b.upload(i, V0, ID0)
    b.run(1, graph=False)
    b.run(STEPS - 1, graph=True)
    for i in CHECK:
        s, V0, ID0 = insts[i]
        Vb, _ = b.download(i)
        bs = b.status(i)
        ost, ot, on, Vo, _ = cases.run_oracle(s, V0, ID0, maxsteps=STEPS, vanalbada_pow=False, nthreads=0)
        assert bs.error == 0 and ost == 0 and bs.steps == on == STEPS
        if prec == M.PREC_FP64:
            assert bs.t == ot
            assert np.array_equal(Vb, Vo), i
            print(f"  instance {i}: bit-identical to the oracle", flush=True)
        else:
            e = rel_err(Vb, Vo)
            print(f"  instance {i}: L1/Linf vs oracle " + ", ".join(f"{k} {a:.2e}/{c:.2e}" for k, (a, c) in e.items()),
                  flush=True)
            assert all(a <= 1e-5 for a, _ in e.values())
    b.close()
