"""Train the learned interface solver on the GPU and store its parameters.

    python m2c_amd/train_surrogate.py [--log2-samples 24] [--epochs 5] [--out m2c_amd/data/sg_riemann_mlp.npy]

Training samples are stiffened-gas Riemann problems solved by the library's exact two-material
solver (the device restatement of ExactRiemannSolverBase::ComputeRiemannSolution); the network is
trained by the library's bf16 MFMA forward/backward kernels with Adam. Writes the FP32 parameters
(.npy, 200963 values) and a JSON report beside them.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import m2c_amd as M  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--log2-samples", type=int, default=24)
    ap.add_argument("--holdout", type=int, default=1 << 16)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=5)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-min", type=float, default=1e-5)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--init-seed", type=int, default=5)
    ap.add_argument("--out", default=M.SURROGATE_PARAMS)
    a = ap.parse_args()
    net = M.Mlp(seed=a.init_seed)
    t0 = time.time()
    rep = net.train(n_samples=1 << a.log2_samples, n_holdout=a.holdout, seed=a.seed, batch=a.batch,
                    epochs=a.epochs, lr=a.lr, lr_min=a.lr_min, newton_max=2, tol=1e-6, verbose=True)
    wall = time.time() - t0
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.save(a.out, net.params())
    report = {"config": vars(a), "report": rep.as_dict(), "wall_s": wall}
    with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
