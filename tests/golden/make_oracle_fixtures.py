"""
Writes oracle energies and gradients for the full-size cluster geometries, so
the GPU tests compare against them without rerunning the CPU oracle:

    (H2O)16 RHF/6-31G*  (water16.xyz)  -> oracle_water16_631gs.json
    (H2O)8  RHF/cc-pVDZ (water8.xyz)   -> oracle_water8_ccpvdz.json

SCF thresholds 1e-11 Eh / 1e-9, Schwarz 1e-12, gradient screening 1e-14,
SAD guess, integral-direct J/K.

    python tests/golden/make_oracle_fixtures.py
"""

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

BOHR = 0.529177210903


def read_xyz(path):
    lines = open(path).read().splitlines()
    n = int(lines[0])
    sym = [l.split()[0] for l in lines[2:2 + n]]
    pos = np.array([[float(v) for v in l.split()[1:4]] for l in lines[2:2 + n]])
    return sym, pos / BOHR


def run(xyz, basis, out):
    sym, x = read_xyz(os.path.join(HERE, xyz))
    t = time.time()
    st, g = O.energy_and_gradient(sym, x, basis, conv_energy=1e-11, conv_commutator=1e-9, stored=False)
    rec = {"geometry": xyz, "basis": basis, "energy": st.energy, "gradient": g.tolist(),
           "iterations": st.iterations, "converged": bool(st.converged), "commutator": st.err,
           "seconds": time.time() - t}
    with open(os.path.join(HERE, out), "w") as f:
        json.dump(rec, f, indent=1)
    print(out, st.energy, st.iterations, rec["seconds"])


if __name__ == "__main__":
    O.set_threads(os.cpu_count())
    run("water8.xyz", "cc-pvdz", "oracle_water8_ccpvdz.json")
    run("water16.xyz", "6-31g*", "oracle_water16_631gs.json")
