"""Writes tests/golden/shock_bubble_384x192x192_digest.json: the oracle's state of the full-size
water/air shock-bubble (384x192x192, RK2, HLLC + van Albada, exact interface Riemann solves, level
set with reinitialisation every 5 steps) after 150 and 300 time steps.

For each checkpoint the fixture holds the SHA-256 of the primitive state V (nz, ny, nx, 5 float64,
negative zeros folded into positive ones), of the material IDs (int32) and of the level set
(float64), and per-field sums and norms. The oracle runs with the van Albada power evaluated as
dq*dq (what the device kernels compute). A second run evaluates it with pow(dq, 2) as
Reconstructor.h:107-108 does; the fixture records how far the two runs drift apart (per-field
relative L1 and L-infinity at both checkpoints): the sensitivity of this flow to last-bit changes of
the reconstruction.

Run from the repository root: python tests/golden/make_shock_bubble_digest.py [--threads N]
(about 10-20 minutes per run on 8 cores).
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from m2c_amd import workloads as W  # noqa: E402
from oracle import pyoracle as po  # noqa: E402
from tests.cases import oracle_config  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shock_bubble_384x192x192_digest.json")
CHECKPOINTS = (150, 300)
FIELDS = ("rho", "u", "v", "w", "p")


def digest(V, ID, phi):
    """SHA-256 and norms of an interior state (the layout m2c_download_state returns)."""
    Vc = np.ascontiguousarray(V, dtype=np.float64) + 0.0       # -0.0 + 0.0 = +0.0
    out = {"sha256_V": hashlib.sha256(Vc.tobytes()).hexdigest(),
           "sha256_ID": hashlib.sha256(np.ascontiguousarray(ID, dtype=np.int32).tobytes()).hexdigest(),
           "sha256_phi": hashlib.sha256((np.ascontiguousarray(phi, dtype=np.float64) + 0.0).tobytes()).hexdigest()}
    out["fields"] = {f: {"sum": float(Vc[..., i].sum()), "l1": float(np.abs(Vc[..., i]).sum()),
                         "linf": float(np.abs(Vc[..., i]).max())} for i, f in enumerate(FIELDS)}
    out["id_counts"] = {str(int(k)): int(v) for k, v in zip(*np.unique(ID, return_counts=True))}
    out["phi_sum"] = float(phi.sum())
    return out


def drift(Va, Vb):
    """Relative L1 and L-infinity per field (velocities against the whole velocity's norms)."""
    vel1 = float(np.abs(Vb[..., 1:4]).sum())
    velinf = float(np.abs(Vb[..., 1:4]).max())
    out = {}
    for i, f in enumerate(FIELDS):
        d = np.abs(Va[..., i] - Vb[..., i])
        n1 = vel1 if 1 <= i <= 3 else float(np.abs(Vb[..., i]).sum())
        ninf = velinf if 1 <= i <= 3 else float(np.abs(Vb[..., i]).max())
        out[f] = {"l1": float(d.sum() / n1), "linf": float(d.max() / ninf)}
    return out


def run(settings, V0, ID0, phi0, pow_mode, threads):
    o = po.Oracle(oracle_config(settings, vanalbada_pow=pow_mode, nthreads=threads))
    V, ID = o.ghosted(V0, ID0)
    phi = o.ghosted_phi(phi0)
    states = {}
    done = 0
    for chk in CHECKPOINTS:
        t0 = time.time()
        st, t, n = o.run(V, ID, 1e300, chk - done, phi=phi)
        done += n
        if st != 0 or done != chk:
            raise RuntimeError(f"oracle stopped at step {done} (status {st})")
        print(f"  pow={pow_mode}: step {done} t={t:.6e} ({time.time() - t0:.0f} s)", flush=True)
        states[chk] = (V[1:-1, 1:-1, 1:-1].copy(), ID[1:-1, 1:-1, 1:-1].copy(), phi[1:-1, 1:-1, 1:-1].copy(), t)
    a, b, c = o.counters()
    o.close()
    return states, {"reverted_faces": a, "clipped_cells": b, "riemann_errors": c}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    settings, V0, ID0, phi0 = W.shock_bubble()
    fx = {"workload": "shock_bubble 384x192x192, RK2, HLLC, van Albada, exact interface Riemann, level set, "
                      "reinitialisation every 5 steps, CFL 0.5, FP64",
          "generator": "tests/golden/make_shock_bubble_digest.py", "checkpoints": {}}
    print("oracle, van Albada dq*dq", flush=True)
    sq, cnt = run(settings, V0, ID0, phi0, False, args.threads)
    for chk, (V, ID, phi, t) in sq.items():
        fx["checkpoints"][str(chk)] = dict(digest(V, ID, phi), t=t)
    fx["counters_dqdq"] = cnt
    with open(OUT, "w") as f:
        json.dump(fx, f, indent=1)
    print("oracle, van Albada pow(dq, 2)", flush=True)
    pw, cnt2 = run(settings, V0, ID0, phi0, True, args.threads)
    fx["pow_mode"] = {str(chk): dict(drift(pw[chk][0], sq[chk][0]),
                                     id_cells_differing=int((pw[chk][1] != sq[chk][1]).sum()),
                                     identical=bool(np.array_equal(pw[chk][0], sq[chk][0])))
                      for chk in CHECKPOINTS}
    fx["counters_pow"] = cnt2
    with open(OUT, "w") as f:
        json.dump(fx, f, indent=1)
    print(json.dumps(fx["pow_mode"], indent=1))


if __name__ == "__main__":
    main()
