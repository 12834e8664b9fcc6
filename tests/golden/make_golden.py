"""Writes the committed golden fixtures under tests/golden/.

grid_reference_hills.npy: the config-1 grid ([0,1], 1000 bins) after depositing the
238 hills of the reference's own test fixture test/test_cvset/HILLS (copied here as
HILLS_reference_test_cvset, data only) with the FP64 oracle's add_to_grid restatement.
The oracle itself is pinned by the closed-form check in tests/test_oracle.py.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as orc  # noqa: E402

rows = np.loadtxt(os.path.join(HERE, "HILLS_reference_test_cvset"), comments="#")
g = orc.Grid([0.0], [1.0], [1000])
for r in rows:
    g.deposit([r[1]], r[2], r[3])
np.save(os.path.join(HERE, "grid_reference_hills.npy"), g.g)
print("wrote grid_reference_hills.npy", g.g.max())
