"""Full-size parity of the headline configuration: the 384x192x192 water/air shock-bubble, FP64,
300 steps on the GPU, checked at steps 150 and 300 against the digest the oracle wrote on the CPU
(tests/golden/shock_bubble_384x192x192_digest.json, made by tests/golden/make_shock_bubble_digest.py):
SHA-256 of the primitive state, the material IDs and the level set (bit-identical), the simulated
time, and per-field sums and norms (which name the first field that differs should the hashes not
match)."""
import json
import os

import numpy as np
import pytest

import m2c_amd as M
from m2c_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "shock_bubble_384x192x192_digest.json")
FIELDS = ("rho", "u", "v", "w", "p")


def test_full_size_shock_bubble_matches_oracle_digest():
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from bench import state_digest
    with open(GOLDEN) as f:
        fx = json.load(f)
    settings, V0, ID0, phi0 = W.shock_bubble()
    print("case shock-bubble 384x192x192 fp64 300 steps", flush=True)
    s = M.Solver(W.config_from(settings))
    s.upload(V0, ID0)
    s.upload_levelset(phi0)
    done = 0
    for chk in (150, 300):
        st = s.run(1e300, chk - done, sync_every=0)
        done = chk
        V, ID = s.download()
        phi = s.download_levelset()
        ref = fx["checkpoints"][str(chk)]
        d = state_digest(V, ID, phi)
        Vc = V + 0.0
        for i, f in enumerate(FIELDS):
            got = {"sum": float(Vc[..., i].sum()), "l1": float(np.abs(Vc[..., i]).sum()),
                   "linf": float(np.abs(Vc[..., i]).max())}
            print(f"  step {chk} {f}: sum {got['sum']!r} (oracle {ref['fields'][f]['sum']!r})", flush=True)
        assert st.error == 0 and st.steps == chk
        assert st.t == ref["t"]
        assert d["sha256_ID"] == ref["sha256_ID"]
        assert d["sha256_V"] == ref["sha256_V"]
        assert d["sha256_phi"] == ref["sha256_phi"]
    s.close()
