"""Pins of MultiPhase.ConstantReconstructionDepth in the oracle (SpaceOperator::TagNodesOutsideConRecDepth,
SpaceOperator.cpp:2861-2920; Reconstructor.cpp:434-447): cells whose level set at t_n lies within
the depth of the interface reconstruct constantly.

* A depth that covers every cell gives the constant-reconstruction run bit for bit, a zero depth the
  linear one, and a depth of a few cells something in between that differs from both;
* the selection is by |phi(t_n)|: a level set shifted far from the cells (|phi| above the depth
  everywhere) leaves the linear run unchanged.
"""
import numpy as np

from m2c_amd import workloads as W
from tests.cases import run_oracle


def sb():
    settings, V0, ID0, phi0 = W.shock_bubble(scale=0.0625, reinit_freq=3)
    return settings, V0, ID0, phi0


def run(settings, V0, ID0, phi0, **over):
    ost, t, n, V, _, ID, phi = run_oracle(dict(settings, **over), V0, ID0, maxsteps=6, phi0=phi0, vanalbada_pow=False)
    assert ost == 0 and n == 6
    return V


def test_depth_limits():
    settings, V0, ID0, phi0 = sb()
    lin = run(settings, V0, ID0, phi0)
    const = run(settings, V0, ID0, phi0, rec_linear=False)
    assert np.array_equal(run(settings, V0, ID0, phi0, conrec_depth=1e30), const)
    assert np.array_equal(run(settings, V0, ID0, phi0, conrec_depth=0.0), lin)
    h = settings["box"][1] / settings["nx"]
    mid = run(settings, V0, ID0, phi0, conrec_depth=2.5 * h)
    assert np.max(np.abs(mid - lin)) > 0 and np.max(np.abs(mid - const)) > 0


def test_depth_selects_by_the_level_set():
    # one residual of a state with slopes at the interface: the depth with the level set shifted far
    # from zero (1e6) selects no cell and changes nothing; unshifted it changes the residual
    from oracle import pyoracle as po
    from tests.cases import oracle_config
    settings, V0, ID0, phi0 = sb()
    V0 = V0.copy()
    nz, ny, nx = ID0.shape
    x = (np.arange(nx) + 0.5) / nx
    V0[..., 0] *= 1.0 + 0.05 * np.sin(2 * np.pi * x)     # slopes next to the bubble's interface
    V0[..., 1] += 0.5 * np.cos(2 * np.pi * x)
    h = settings["box"][1] / settings["nx"]
    out = []
    for depth, shift in ((0.0, 0.0), (2.5 * h, 1e6), (2.5 * h, 0.0)):
        o = po.Oracle(oracle_config(dict(settings, conrec_depth=depth), vanalbada_pow=False))
        V, ID = o.ghosted(V0, ID0)
        phi = o.ghosted_phi(phi0 + shift)
        st, R = o.residual(V, ID, phi=phi)
        assert st == 0
        out.append(R)
    assert np.array_equal(out[0], out[1])
    assert not np.array_equal(out[0], out[2])
