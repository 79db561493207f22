"""
Writes the water-cluster geometries the engine is tested and timed on.

Recipe: 16 / 8 / 32 copies of the water molecule O (0, 0, 0.1173) A,
H (0, +-0.7572, -0.4692) A, each rotated about its O by a uniformly random
rotation (scipy Rotation.random driven by numpy default_rng(seed)) and placed
with O uniform in an L-A cube, rejecting a placement until every O-O
distance is >= 2.7 A. Seed 1 for all three. Output: XYZ files (angstrom).

    python tests/golden/make_clusters.py
"""

import os

import numpy as np
from scipy.spatial.transform import Rotation

HERE = os.path.dirname(os.path.abspath(__file__))
WATER = np.array([[0.0, 0.0, 0.1173], [0.0, 0.7572, -0.4692], [0.0, -0.7572, -0.4692]])


def cluster(nmol: int, box: float, seed: int = 1, dmin: float = 2.7) -> np.ndarray:
    rng = np.random.default_rng(seed)
    oxy, mols = [], []
    while len(mols) < nmol:
        rot = Rotation.random(random_state=rng)
        o = rng.uniform(0.0, box, 3)
        if any(np.linalg.norm(o - p) < dmin for p in oxy):
            continue
        oxy.append(o)
        mols.append(rot.apply(WATER - WATER[0]) + o)
    return np.concatenate(mols)


def write_xyz(path: str, pos: np.ndarray, comment: str) -> None:
    n = len(pos)
    sym = ["O", "H", "H"] * (n // 3)
    with open(path, "w") as f:
        f.write(f"{n}\n{comment}\n")
        for s, x in zip(sym, pos):
            f.write(f"{s} {x[0]:.10f} {x[1]:.10f} {x[2]:.10f}\n")


if __name__ == "__main__":
    for n, box in ((16, 8.0), (8, 6.5), (32, 10.0)):
        write_xyz(os.path.join(HERE, f"water{n}.xyz"), cluster(n, box),
                  f"(H2O){n} in a {box} A cube, seed 1, O-O >= 2.7 A")
