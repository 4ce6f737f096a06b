"""Writes the small fixtures under tests/golden/.

tfim_couplings.json: the seeded J_ij of the transverse-field spin-glass workloads (seed = workload
index 0-4, spin counts 3, 6, 8, 4, 4) from the oracle's SplitMix64 + Box-Muller stream; the product's
C++ set-up must reproduce them bit for bit.

legacy_outputEnergy.txt is copied verbatim from the reference's examples/outputEnergy.txt
(MIT licence, Copyright (c) 2015 Masashi Tsuchimoto and Yoshitaka Tanimura): the golden FP32
energy trace of the legacy spin glass, 1981 rows "time energy".
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def main():
    out = {"generator": "SplitMix64 -> 53-bit uniform -> Box-Muller (cos branch), J[i][j] for i<j row by row",
           "couplings": {}}
    for seed, spins in [(0, 3), (1, 6), (2, 8), (3, 4), (4, 4)]:
        J = O.tfim_couplings(spins, seed)
        out["couplings"][f"seed{seed}"] = {"seed": seed, "spins": spins, "J": J.tolist()}
    with open(os.path.join(HERE, "tfim_couplings.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
