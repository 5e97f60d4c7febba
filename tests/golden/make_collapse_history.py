"""Writes collapse_history.npz: the FP64 oracle's history of the 4,096-ball
granular collapse (scenes.granular_collapse) over its full 10,000 steps,
sampled every 100 steps: translational and rotational kinetic energy, mean
and maximum ball height.  The oracle is the checker; this fixture lets a GPU
test compare the whole run without recomputing 41 million FP64 ball-steps.

Run from the repository root: python tests/golden/make_collapse_history.py
"""
import math
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from this_studio_amd import scenes  # noqa: E402


def history(o, r: float, rho: float):
    m = 4.0 / 3.0 * math.pi * r ** 3 * rho
    v = o.get("vel")
    w = o.get("angVel")
    z = o.get("pos")[:, 2]
    return [0.5 * m * float((v * v).sum()), 0.5 * 0.4 * m * r * r * float((w * w).sum()), float(z.mean()),
            float(z.max())]


def main():
    s = scenes.granular_collapse()
    o = oracle.OracleDEM(s)
    rows = [[0] + history(o, 0.005, 2500.0)]
    t0 = time.time()
    for k in range(1, 101):
        o.step(100)
        rows.append([100 * k] + history(o, 0.005, 2500.0))
    print(f"10,000 oracle steps in {time.time() - t0:.1f} s", file=sys.stderr)
    out = Path(__file__).resolve().parent / "collapse_history.npz"
    np.savez_compressed(out, history=np.array(rows),
                        columns=np.array(["step", "ke_trans", "ke_rot", "z_mean", "z_max"]))


if __name__ == "__main__":
    main()
