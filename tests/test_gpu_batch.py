"""Batched instances (one persistent launch per RK stage over per-XCD tile queues, a hipGraph per
step, autotuned launch shape): every instance must evolve bit for bit as it does in its own
context, with and without the graph, before and after autotuning, in FP64 and mixed precision, and
with the Roe and LLF fluxes in FP64."""
import numpy as np
import pytest

import m2c_amd as M
from m2c_amd import workloads as W

pytestmark = pytest.mark.gpu


def solo(settings, V0, ID0, steps, **over):
    s = M.Solver(W.config_from(settings, **over))
    s.upload(V0, ID0)
    st = s.run(1e300, steps, sync_every=0)
    V, _ = s.download()
    s.close()
    return st, V


@pytest.mark.parametrize("prec", [M.PREC_FP64, M.PREC_MIXED])
@pytest.mark.parametrize("ts", [M.TS_RK2, M.TS_RK3])
def test_batch_matches_single_instances(prec, ts):
    insts = W.batched_riemann(n=6, lo=6, hi=22, seed=7)
    cfgs = [W.config_from(s, precision=prec, ts=ts) for s, _, _ in insts]
    print(f"batch of {len(insts)} (prec {prec}, ts {ts}): " + ", ".join(f"{s['nx']}x{s['ny']}x{s['nz']}" for s, _, _ in insts),
          flush=True)
    b = M.Batch(cfgs)
    for i, (s, V0, ID0) in enumerate(insts):
        b.upload(i, V0, ID0)
    b.run(4, graph=False)
    b.run(5, graph=True)
    for i, (s, V0, ID0) in enumerate(insts):
        st, V = solo(s, V0, ID0, 9, precision=prec, ts=ts)
        Vb, _ = b.download(i)
        bs = b.status(i)
        assert bs.error == 0 and bs.steps == st.steps == 9
        assert bs.t == st.t
        assert np.array_equal(Vb, V), i
    b.close()


@pytest.mark.parametrize("flux", [M.FLUX_ROE, M.FLUX_LLF])
def test_batch_roe_and_llf_match_single_instances(flux):
    insts = W.batched_riemann(n=4, lo=6, hi=18, seed=5)
    cfgs = [W.config_from(s, flux=flux, ts=M.TS_RK3) for s, _, _ in insts]
    print(f"batch of {len(insts)} (flux {flux})", flush=True)
    b = M.Batch(cfgs)
    for i, (s, V0, ID0) in enumerate(insts):
        b.upload(i, V0, ID0)
    b.run(3, graph=False)
    b.run(3, graph=True)
    for i, (s, V0, ID0) in enumerate(insts):
        st, V = solo(s, V0, ID0, 6, flux=flux, ts=M.TS_RK3)
        Vb, _ = b.download(i)
        bs = b.status(i)
        assert bs.error == 0 and bs.steps == st.steps == 6 and bs.t == st.t
        assert np.array_equal(Vb, V), i
    b.close()


def test_batch_autotune_restores_state_and_keeps_results():
    insts = W.batched_riemann(n=5, lo=8, hi=30, seed=11)
    cfgs = [W.config_from(s, precision=M.PREC_MIXED) for s, _, _ in insts]
    b = M.Batch(cfgs)
    for i, (s, V0, ID0) in enumerate(insts):
        b.upload(i, V0, ID0)
    b.run(2)
    before = [b.download(i)[0] for i in range(len(insts))]
    tune = b.autotune(trial_steps=2)
    print(f"autotune: kchunk {tune.kchunk} blocks {tune.blocks} {tune.ms_per_step:.3f} ms/step over "
          f"{tune.candidates} candidates", flush=True)
    assert tune.candidates >= 6
    after = [b.download(i)[0] for i in range(len(insts))]
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    assert all(b.status(i).steps == 2 for i in range(len(insts)))
    b.run(3)
    for i, (s, V0, ID0) in enumerate(insts):
        st, V = solo(s, V0, ID0, 5, precision=M.PREC_MIXED)
        assert np.array_equal(b.download(i)[0], V)


@pytest.mark.parametrize("prec", [M.PREC_FP64, M.PREC_MIXED])
def test_batch_tile_shapes_match_single_instances(prec):
    # the 16 x 16 column tiles (partial tiles at every instance's edges: edges 9..37) and the 32 x 8 ones
    insts = W.batched_riemann(n=5, lo=9, hi=37, seed=13, two_every=3, reinit_freq=2)
    b = M.Batch([W.config_from(x[0], precision=prec) for x in insts])
    for i, x in enumerate(insts):
        b.upload(i, x[1], x[2])
        if len(x) == 4:
            b.upload_levelset(i, x[3])
    b.set_tile(16, 16)
    b.run(2, graph=False)
    b.set_tile(32, 8)
    b.run(2, graph=True)
    b.set_tile(16, 16)
    b.set_tuning(4, 256)
    b.run(2, graph=True)
    with pytest.raises(M.M2CError):
        b.set_tile(64, 4)
    for i, x in enumerate(insts):
        s = M.Solver(W.config_from(x[0], precision=prec))
        s.upload(x[1], x[2])
        if len(x) == 4:
            s.upload_levelset(x[3])
        st = s.run(1e300, 6, sync_every=0)
        V, ID = s.download()
        phi = s.download_levelset() if len(x) == 4 else None
        s.close()
        Vb, IDb = b.download(i)
        assert b.status(i).steps == st.steps == 6 and b.status(i).t == st.t, i
        assert np.array_equal(Vb, V) and np.array_equal(IDb, ID), i
        if phi is not None:
            assert np.array_equal(b.download_levelset(i), phi), i
    b.close()
