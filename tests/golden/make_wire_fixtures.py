"""Writes the byte-string format fixtures (format stability vectors):

  wire_hodlr_n64_l2_r2.pphl   HODLR n=64, 2 levels, rank 2, with the arena
                              arena[i] = 0.25 * i - 7 (exact in binary64)
  wire_factor_prog_w2.pphl    factor program of that matrix on 2 workers

Run from the repository root: python tests/golden/make_wire_fixtures.py
"""
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
pp = importlib.import_module("pastel-palettes_amd")


def fixture_matrix():
    hm = pp.HMatrix("hodlr", 64, levels=2, rank=2)
    arena = 0.25 * np.arange(hm.storage, dtype=np.float64) - 7.0
    return hm, arena


def main():
    hm, arena = fixture_matrix()
    with open(os.path.join(HERE, "wire_hodlr_n64_l2_r2.pphl"), "wb") as f:
        f.write(hm.to_bytes(arena))
    with open(os.path.join(HERE, "wire_factor_prog_w2.pphl"), "wb") as f:
        f.write(pp.Program.factor(hm, 2).to_bytes())


if __name__ == "__main__":
    main()
