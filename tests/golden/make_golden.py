"""Writes tests/golden/carsten_potential.json: outputs of the reference's own code-carsten
potential_energy() / LennardJones() (code-carsten/potential_energy.hh:6-15, lennardjones.hh:5-9),
compiled from the reference sources by `make -C oracle ref` into oracle/_ref/.

Inputs are seeded clusters (numpy default_rng) in a few units of the default sigma = 0.1; only
inputs and outputs are stored.  Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402


def main():
    oracle.build()
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libref_carsten.so not built: the reference tree is needed once")
    cases = []
    for n, seed, scale in [(2, 1, 0.15), (10, 2, 0.4), (64, 3, 0.8), (200, 4, 1.5)]:
        rng = np.random.default_rng(seed)
        x = rng.uniform(0.0, scale, size=(n, 3))
        v = ref.ref_potential_energy(np.ascontiguousarray(x), n)
        cases.append({"n": n, "seed": seed, "x": x.tolist(), "potential_energy": v})
    pair = []
    for r in [0.09, 0.1, 0.1122462048309373, 0.15, 0.25, 1.0]:
        pair.append({"r": r, "n": 12, "a": 1e-12, "b": 1e-6, "value": ref.ref_lennard_jones(r, 12, 1e-12, 1e-6)})
        pair.append({"r": r * 10, "n": 12, "a": 4.0, "b": 4.0, "value": ref.ref_lennard_jones(r * 10, 12, 4.0, 4.0)})
    out = {"source": "code-carsten potential_energy.hh / lennardjones.hh compiled by oracle/Makefile ref",
           "cases": cases, "lennard_jones": pair}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "carsten_potential.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
