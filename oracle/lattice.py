"""Lattice tables for D2Q8 (elasticity) and D2Q9 (Darcy).  Oracle / test use only.

Restates ``src/lattice.py``: velocity order ``:124-131``, weights ``:161-168``,
opposite/main/right/left construction ``:51-109``.
"""
import numpy as np


class OracleLattice:
    def __init__(self, name, cx, cy, w):
        self.name = name
        self.d = 2
        self.q = len(cx)
        # (d, q) like the reference's ``lattice.c``
        self.c = np.array([cx, cy], dtype=np.int64)
        self.w = np.array(w, dtype=np.float64)
        ct = self.c.T
        # lattice.py:51-65: index of -c_i
        self.opp_indices = np.array(
            [[tuple(v) for v in ct.tolist()].index(tuple((-ct[i]).tolist())) for i in range(self.q)])
        # lattice.py:81-109
        self.main_indices = np.nonzero(np.abs(ct[:, 0]) + np.abs(ct[:, 1]) == 1)[0]
        self.right_indices = np.nonzero(ct[:, 0] == 1)[0]
        self.left_indices = np.nonzero(ct[:, 0] == -1)[0]

    def __str__(self):
        return self.name


def D2Q8():
    """lattice.py:124-127 (velocities), :161-163 (weights 1/9 axis, 1/36 diagonal)."""
    return OracleLattice("D2Q8",
                         [1, 0, -1, 0, 1, -1, -1, 1],
                         [0, 1, 0, -1, 1, 1, -1, -1],
                         [1 / 9] * 4 + [1 / 36] * 4)


def D2Q9():
    """lattice.py:128-131 (velocities), :164-166 (weights 4/9, 1/9, 1/36)."""
    cx = [0, 0, 0, 1, -1, 1, -1, 1, -1]
    cy = [0, 1, -1, 0, 1, -1, 0, 1, -1]
    w = []
    for x, y in zip(cx, cy):
        n = abs(x) + abs(y)
        w.append(4 / 9 if n == 0 else (1 / 9 if n == 1 else 1 / 36))
    return OracleLattice("D2Q9", cx, cy, w)
