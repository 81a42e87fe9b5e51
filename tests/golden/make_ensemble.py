"""Writes ensemble_64.json: the seeded table of the 64-member Riemann-strip
ensemble (Nx drawn from {100, 200, 400, 800, 1600}; left/right primitive
states rho ~ U[0.1, 2], u ~ U[-1, 1], p ~ U[0.1, 10], v = 0; interface at
x = 0.5).  The doubles are committed so that every run sees the same members.

    python tests/golden/make_ensemble.py
"""
import json
from pathlib import Path

import numpy as np

SEED = 20261203
N_MEMBERS = 64
NX_CHOICES = (100, 200, 400, 800, 1600)


def draw(seed: int = SEED, n: int = N_MEMBERS) -> list[dict]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx = int(rng.choice(NX_CHOICES))
        left = [float(rng.uniform(0.1, 2.0)), float(rng.uniform(-1.0, 1.0)), 0.0,
                float(rng.uniform(0.1, 10.0))]
        right = [float(rng.uniform(0.1, 2.0)), float(rng.uniform(-1.0, 1.0)), 0.0,
                 float(rng.uniform(0.1, 10.0))]
        out.append({"nx": nx, "left": left, "right": right})
    return out


if __name__ == "__main__":
    path = Path(__file__).with_name("ensemble_64.json")
    path.write_text(json.dumps({"seed": SEED, "ny": 4, "x0": 0.5, "gamma": 1.4,
                                "members": draw()}, indent=1) + "\n")
    print(path)
