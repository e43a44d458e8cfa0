"""m2c_riemann_batch_noalloc (no device allocation; the solver's path cache in the library's fixed
device array) against m2c_riemann_batch on the same problems: two calls in a row on one context, the
second with fewer problems, each bit-identical to the allocating entry point. Inputs and outputs are
device buffers the caller owns (torch tensors here)."""
import ctypes as C

import numpy as np
import pytest
import torch

import m2c_amd as M
from m2c_amd import workloads as W

pytestmark = pytest.mark.gpu


def problems(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    idm = rng.integers(0, 2, n).astype(np.int32)
    idp = rng.integers(0, 2, n).astype(np.int32)
    Vm = np.zeros((n, 5))
    Vp = np.zeros((n, 5))
    for V, ids in ((Vm, idm), (Vp, idp)):
        V[:, 0] = np.where(ids == 0, rng.uniform(990.0, 1100.0, n), rng.uniform(0.5, 2.0, n))
        V[:, 1:4] = rng.uniform(-60.0, 60.0, (n, 3))
        V[:, 4] = np.where(ids == 0, rng.uniform(1e5, 1e8, n), rng.uniform(5e4, 3e5, n))
    return d, Vm, idm, Vp, idp


def test_noalloc_matches_allocating_entry_point():
    settings, _, _, _ = W.shock_bubble(scale=0.125)
    s = M.Solver(W.config_from(settings))
    L = s.L
    L.m2c_riemann_batch_noalloc.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 10
    dev = torch.device("cuda:0")
    for call, n in enumerate((3000, 700)):
        print(f"case noalloc call {call} n={n}", flush=True)
        d, Vm, idm, Vp, idp = problems(n, 40 + call)
        st, Vs, mid, Vsm, Vsp = s.riemann(d, Vm, idm, Vp, idp)
        t = {k: torch.from_numpy(np.ascontiguousarray(a)).to(dev) for k, a in
             (("d", d), ("Vm", Vm), ("idm", idm), ("Vp", Vp), ("idp", idp))}
        oVs = torch.full((n, 5), np.nan, dtype=torch.float64, device=dev)
        oVsm = torch.full((n, 5), np.nan, dtype=torch.float64, device=dev)
        oVsp = torch.full((n, 5), np.nan, dtype=torch.float64, device=dev)
        oid = torch.full((n,), -7, dtype=torch.int32, device=dev)
        ost = torch.full((n,), -7, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        rc = L.m2c_riemann_batch_noalloc(s.h, n, t["d"].data_ptr(), t["Vm"].data_ptr(), t["idm"].data_ptr(),
                                         t["Vp"].data_ptr(), t["idp"].data_ptr(), oVs.data_ptr(), oid.data_ptr(),
                                         oVsm.data_ptr(), oVsp.data_ptr(), ost.data_ptr())
        assert rc == 0
        torch.cuda.synchronize()
        assert np.array_equal(ost.cpu().numpy(), st) and np.array_equal(oid.cpu().numpy(), mid)
        assert np.array_equal(oVs.cpu().numpy(), Vs)
        assert np.array_equal(oVsm.cpu().numpy(), Vsm)
        assert np.array_equal(oVsp.cpu().numpy(), Vsp)
        print(f"  bit-identical, {int((st != 0).sum())} solver failures", flush=True)
    s.close()
