#!/usr/bin/env python3
"""Writes tests/golden/force_vectors.json from the reference's turbulence driver
(turb_driver_mod.f90), built by oracle/ref/build_ref.py with drive_turbulence (needs
/root/reference; run from the repository root):
  taus   rand_taus streams (turb_driver_mod.f90:104-126) for recorded seed triples
  idft   direct_idft_1d (turb_driver_mod.f90:461-486) on seeded inputs in(-k_max:k_max-1),
         k_max = 2, over the padded cell range of a 16-cell axis (ilow = -5, ihi = 22)
Floats are stored as Python reprs, which round-trip exactly.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from refutil import Reference  # noqa: E402

TAUS_SEEDS = [(1234, 5678, 9123), (2, 8, 16), (-123456789, 987654321, -5)]
IDFT_SEEDS = [901, 902]


def main():
    R = Reference("vl_hlld_p16_drive")
    out = {"taus": [], "idft": []}
    for s in TAUS_SEEDS:
        out["taus"].append({"seeds": list(s), "values": [float(v) for v in R.rand_taus(s, 64)]})
    for seed in IDFT_SEEDS:
        rng = np.random.default_rng(seed)
        a, b = rng.standard_normal(4), rng.standard_normal(4)
        ore, oim = R.idft_1d(a, b, -5, 22, 16)
        out["idft"].append({"seed": seed, "in_re": [float(v) for v in a], "in_im": [float(v) for v in b],
                            "ilow": -5, "ihi": 22, "nnx": 16, "k_max": 2,
                            "out_re": [float(v) for v in ore], "out_im": [float(v) for v in oim]})
    path = os.path.join(HERE, "force_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", path)


if __name__ == "__main__":
    main()
