"""Load the committed reference meshes (tests/golden/*.npz) as problem dictionaries."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import scipy.sparse as sp

GOLDEN = Path(__file__).resolve().parent / "golden"


def _load(name: str):
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    d = {k: z[k] for k in z.files}
    shape = tuple(int(v) for v in d.pop("adjacency_shape"))
    row = d.pop("adjacency_row")
    col = d.pop("adjacency_col")
    adj = sp.coo_matrix((np.ones(row.size), (row, col)), shape=shape)
    for k, v in list(d.items()):
        if v.shape == ():
            d[k] = v[()]
    return d, adj


def plaplace(level: int):
    d, adj = _load(f"plaplace_l{level}")
    d["elems"] = d["elems"].astype(np.int64)
    d["freedofs"] = d["freedofs"].astype(np.int64)
    d["p"] = float(d["p"])
    return d, adj


def gl(level: int):
    d, adj = _load(f"gl_l{level}")
    d["elems"] = d["elems"].astype(np.int64)
    d["freedofs"] = d["freedofs"].astype(np.int64)
    d["eps"] = float(d["eps"])
    return d, adj


def he(level: int = 1):
    """HE mesh in the assembler's dictionary form (hyperelasticity/support/mesh.py:40-82)."""
    raw, adj = _load(f"he_l{level}")
    return raw, adj


def published_energies() -> dict:
    return json.loads((GOLDEN / "published_energies.json").read_text())
