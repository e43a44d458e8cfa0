"""Two-material instances in a batch: gas/gas Riemann problems whose split plane is a material
interface (exact Riemann solves at the interface faces, a level set advected and reinitialised,
material IDs and swept-cell states updated) advanced beside single-material instances in the
batch's persistent stage launches. Every instance must evolve bit for bit as in its own context
(state, material IDs, level set), with and without the step graphs and across reinitialisation
steps, in FP64 and in the mixed form; a full-size two-material instance of the 64-instance workload
must be bit-identical to the oracle; autotuning must leave the two-material state untouched."""
import numpy as np
import pytest

import m2c_amd as M
from m2c_amd import workloads as W
from tests import cases

pytestmark = pytest.mark.gpu


def solo(inst, steps, **over):
    settings, V0, ID0 = inst[:3]
    s = M.Solver(W.config_from(settings, **over))
    s.upload(V0, ID0)
    if len(inst) == 4:
        s.upload_levelset(inst[3])
    st = s.run(1e300, steps, sync_every=0)
    V, ID = s.download()
    phi = s.download_levelset() if len(inst) == 4 else None
    s.close()
    return st, V, ID, phi


def make_batch(insts, **over):
    b = M.Batch([W.config_from(x[0], **over) for x in insts])
    for i, x in enumerate(insts):
        b.upload(i, x[1], x[2])
        if len(x) == 4:
            b.upload_levelset(i, x[3])
    return b


@pytest.mark.parametrize("prec", [M.PREC_FP64, M.PREC_MIXED])
def test_two_material_batch_matches_single_contexts(prec):
    insts = W.batched_riemann(n=6, lo=8, hi=24, seed=7, two_every=2, reinit_freq=3)
    print(f"two-material batch (prec {prec}): " + ", ".join(
        f"{x[0]['nx']}x{x[0]['ny']}x{x[0]['nz']}{'*' if len(x) == 4 else ''}" for x in insts), flush=True)
    assert sum(len(x) == 4 for x in insts) == 3
    b = make_batch(insts, precision=prec)
    b.run(4, graph=False)   # reinitialisation at step 3
    b.run(5, graph=True)    # ... and at steps 6 and 9, replayed from the graphs
    for i, x in enumerate(insts):
        st, V, ID, phi = solo(x, 9, precision=prec)
        Vb, IDb = b.download(i)
        bs = b.status(i)
        assert bs.error == 0 and st.error == 0 and bs.steps == st.steps == 9, i
        assert bs.t == st.t, i
        assert np.array_equal(Vb, V), i
        assert np.array_equal(IDb, ID), i
        if phi is not None:
            assert np.array_equal(b.download_levelset(i), phi), i
            assert bs.reinit_iterations == st.reinit_iterations > 0, i
    b.close()


def test_two_material_batch_instance_against_oracle():
    # instance 3 of the 64-instance workload with every 4th instance two-material, at full size,
    # beside a single-material one
    insts = W.batched_riemann(n=64, select=[2, 3], two_every=4, reinit_freq=3)
    assert len(insts[1]) == 4
    settings, V0, ID0, phi0 = insts[1]
    print(f"case two-material instance {settings['nx']}x{settings['ny']}x{settings['nz']}", flush=True)
    b = make_batch(insts)
    b.run(2, graph=False)
    b.run(2, graph=True)
    Vb, IDb = b.download(1)
    phib = b.download_levelset(1)
    bs = b.status(1)
    b.close()
    ost, ot, on, Vo, _, IDo, phio = cases.run_oracle(settings, V0, ID0, maxsteps=4, phi0=phi0, vanalbada_pow=False)
    assert bs.error == 0 and ost == 0 and bs.steps == on == 4
    assert bs.t == ot
    assert np.array_equal(IDb, IDo)
    assert np.array_equal(phib, phio)
    assert np.array_equal(Vb, Vo)
    print(f"  bit-identical to the oracle after 4 steps ({int((IDo == 1).sum())} cells of the second gas)",
          flush=True)


def test_two_material_batch_autotune_restores_state():
    insts = W.batched_riemann(n=4, lo=8, hi=20, seed=3, two_every=2, reinit_freq=2)
    b = make_batch(insts)
    b.run(2)
    before = [(b.download(i), b.download_levelset(i) if len(x) == 4 else None) for i, x in enumerate(insts)]
    tune = b.autotune(trial_steps=2)
    print(f"autotune: kchunk {tune.kchunk} blocks {tune.blocks} {tune.ms_per_step:.3f} ms/step", flush=True)
    for i, x in enumerate(insts):
        (V0, ID0), phi0 = before[i]
        V1, ID1 = b.download(i)
        assert np.array_equal(V0, V1) and np.array_equal(ID0, ID1), i
        if phi0 is not None:
            assert np.array_equal(phi0, b.download_levelset(i)), i
        assert b.status(i).steps == 2
    b.run(3)
    for i, x in enumerate(insts):
        st, V, ID, phi = solo(x, 5)
        Vb, IDb = b.download(i)
        assert np.array_equal(Vb, V) and np.array_equal(IDb, ID), i
        if phi is not None:
            assert np.array_equal(b.download_levelset(i), phi), i
    b.close()
