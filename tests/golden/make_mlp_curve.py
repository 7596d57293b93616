"""Writes tests/golden/mlp_loss_curve.json: the FP64 reference's full-batch
Adam trajectory (tests/mlp_ref.py, bf16 rounding emulated at the GPU's
rounding points) on a fixed training set, with its optimal-root labels from
the oracle DP.

Run from the repo root: python tests/golden/make_mlp_curve.py
"""
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import numpy as np  # noqa: E402

import mlp_ref  # noqa: E402
import oracle_lib  # noqa: E402

INIT_SEED = 5
STEPS = 8
LR = 2e-3


def main():
    o = oracle_lib.Oracle()
    gs = mlp_ref.curve_dataset()
    targets = [mlp_ref.optimal_root_set(o, g) for g in gs]
    w0 = mlp_ref.init_weights(INIT_SEED).astype(np.float32).astype(np.float64)
    data = [(g.n, np.array(g.adj + [0] * (64 - g.n), np.uint64), t) for g, t in zip(gs, targets)]
    curve, w = mlp_ref.adam_train(w0, data, STEPS, LR)
    out = {"init_seed": INIT_SEED, "steps": STEPS, "lr": LR,
           "graphs": [{"name": g.name, "n": g.n, "adj": [str(a) for a in g.adj]} for g in gs],
           "targets": [str(t) for t in targets],
           "loss_curve": curve,
           "final_weight_norm": float(np.linalg.norm(w))}
    path = Path(__file__).resolve().parent / "mlp_loss_curve.json"
    path.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", path, curve)


if __name__ == "__main__":
    main()
