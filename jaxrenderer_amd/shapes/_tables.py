"""Loads the mesh constant tables (written by tests/golden/extract_tables.py
from the reference's renderer/shapes/cube.py and capsule.py)."""
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mesh_tables.npz")
with np.load(_PATH, allow_pickle=False) as _z:
    TABLES = {k: _z[k] for k in _z.files}
