"""Pins of the oracle's linear reconstruction at material interfaces (MultiPhase.ReconstructionAtInterface
= Linear, SpaceOperator.cpp:2504-2507, 2529-2535): the exact Riemann problem and the two one-sided
fluxes of an interface face take the reconstructed face states (vr of the minus cell, vl of the plus
cell) instead of the cell states.

* Each material uniform: every limited slope next to the interface vanishes (one of its two
  differences is zero), so the face states are the cell states and the residual equals the
  constant-reconstruction residual bit for bit;
* a smooth field inside each material: the residuals differ at the interface faces only, and the
  difference shrinks with the slopes (halving the perturbation roughly halves it).
"""
import numpy as np

from oracle import pyoracle as po

N = 12


def setup(recon, V, ID):
    cfg = po.make_config(N, N, 4, (0, 1, 0, 1, 0, 1.0 / 3.0), [po.BC_SYMMETRY] * 6,
                         [po.sg(1.4, 0.0), po.sg(1.67, 0.0)], mp_recon=recon)
    o = po.Oracle(cfg)
    Vg, IDg = o.ghosted(V, ID)
    st, R = o.residual(Vg, IDg)
    assert st == 0
    return R[1:-1, 1:-1, 1:-1]


def field(amp):
    c = (np.arange(N) + 0.5) / N
    z, y, x = np.meshgrid((np.arange(4) + 0.5) / 12, c, c, indexing="ij")
    ID = (x + 0.3 * y > 0.55).astype(np.int32)
    V = np.zeros((4, N, N, 5))
    V[..., 0] = np.where(ID == 1, 0.5, 1.0) * (1 + amp * np.sin(2 * np.pi * y))
    V[..., 1] = 0.2 + amp * np.cos(2 * np.pi * x)
    V[..., 2] = -0.1
    V[..., 4] = np.where(ID == 1, 2.0, 1.0) * (1 + amp * np.cos(2 * np.pi * (x + y)))
    return V, ID


def test_uniform_materials_reduce_to_constant_reconstruction():
    V, ID = field(0.0)
    assert np.array_equal(setup(1, V, ID), setup(0, V, ID))


def test_smooth_materials_differ_at_the_interface_only():
    d = []
    for amp in (0.04, 0.02):
        V, ID = field(amp)
        diff = np.abs(setup(1, V, ID) - setup(0, V, ID)).max(axis=-1)
        # cells away from the interface (all six neighbours of their own material) are unchanged
        same = np.ones_like(ID, dtype=bool)
        for ax in (1, 2):
            for s in (1, -1):
                nb = np.roll(ID, s, axis=ax)
                same &= nb == ID
        same[:, 0, :] = same[:, -1, :] = False
        same[:, :, 0] = same[:, :, -1] = False
        assert np.all(diff[same] == 0.0)
        d.append(diff[~same].max())
    assert d[0] > 0 and 1.5 < d[0] / d[1] < 2.6, d
