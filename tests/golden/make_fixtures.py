"""Writes the golden fixtures under tests/golden/ from data files the reference ships.

    python tests/golden/make_fixtures.py

The reference's examples hold two squarer .dm tables (an electron-proton
Coulomb pair density matrix).  They are parsed with pimcplusplus_amd.tables.read_dm,
which follows the reference's own reader, and saved compactly so the tests can
use them where the reference tree is absent.  Only data is stored.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from pimcplusplus_amd.tables import read_dm  # noqa: E402

SOURCES = {
    "testcoul_fit2.npz": "/root/reference/PIMC++/examples/testcoul.dm",
    "ep_beta1_fit2.npz": "/root/reference/PIMC++/examples/ep_beta1.0.dm",
}


def main():
    for out, src in SOURCES.items():
        t = read_dm(src)
        np.savez_compressed(os.path.join(HERE, out), y=t.y, dy=t.dy, grid_type=t.grid_type,
                            grid_start=t.grid_start, grid_end=t.grid_end, smallest_tau=t.smallest_tau,
                            num_terms=t.num_terms, source=os.path.basename(src))
        print(out, t.y.shape)


if __name__ == "__main__":
    main()
