"""Loads the mesh constant tables (see extract_tables.py for their origin)."""
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mesh_tables.npz")
with np.load(_PATH, allow_pickle=False) as _z:
    TABLES = {k: _z[k] for k in _z.files}
