"""Shared helpers of the test-suite: fixtures loading and problem builders."""

from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


class IdxGeom:
    """Minimal duck-typed geometry (idx, n_elements, element_index)."""

    def __init__(self, idx):
        self.idx = [tuple(int(v) for v in t) for t in np.asarray(idx)]
        self.n_elements = len(self.idx)

    def element_index(self, k):
        return self.idx[k]


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def footprints_from(cells, counts, w):
    """Footprint list [(index tuples, weights)] from a fixture's flat arrays."""
    out, o = [], 0
    for c in np.asarray(counts):
        c = int(c)
        out.append(([tuple(int(v) for v in t) for t in np.asarray(cells)[o:o + c]],
                    np.asarray(w)[o:o + c]))
        o += c
    return out
