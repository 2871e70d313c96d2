"""Lattice descriptors D2Q8 (elasticity) and D2Q9 (Darcy).

Same names, constructor and attributes as the reference ``src/lattice.py``
(``Lattice`` :6-213, ``LatticeD2Q8`` :215-234, ``LatticeD2Q9`` :236-258); the
arrays are NumPy (host-side tables; the kernels carry the velocities as
compile-time constants in the same order).  The 3-D lattices are not part of
this path.
"""
import re

import numpy as np

_VELOCITIES = {
    # lattice.py:124-131
    "D2Q8": ([1, 0, -1, 0, 1, -1, -1, 1], [0, 1, 0, -1, 1, 1, -1, -1]),
    "D2Q9": ([0, 0, 0, 1, -1, 1, -1, 1, -1], [0, 1, -1, 0, 1, -1, 0, 1, -1]),
}

_PRECISIONS = {"f32/f32": np.float32, "f32/f16": np.float32, "f64/f64": np.float64, "f64/f32": np.float64,
               "f64/f16": np.float64, "f16/f16": np.float16}


class Lattice(object):
    def __init__(self, name, precision="f32/f32"):
        if name not in _VELOCITIES:
            raise ValueError("Supported lattice types on this path are D2Q8 and D2Q9")
        if precision not in _PRECISIONS:
            raise ValueError("precision not supported")
        self.name = name
        self.precision = precision
        dq = re.findall(r"\d+", name)
        self.d = int(dq[0])
        self.q = int(dq[1])
        self.precisionPolicy = _PRECISIONS[precision]
        cx, cy = _VELOCITIES[name]
        self.c = np.array([cx, cy], dtype=np.int8)  # (d, q)
        self.w = self.construct_lattice_weight().astype(self.precisionPolicy)
        self.cc = self.construct_lattice_moment().astype(self.precisionPolicy)
        self.opp_indices = self.construct_opposite_indices().astype(np.int8)
        self.main_indices = self.construct_main_indices().astype(np.int8)
        self.right_indices = self.construct_right_indices().astype(np.int8)
        self.left_indices = self.construct_left_indices().astype(np.int8)

    def construct_opposite_indices(self):
        ct = self.c.T.astype(np.int64)
        rows = [tuple(r) for r in ct.tolist()]
        return np.array([rows.index(tuple((-ct[i]).tolist())) for i in range(self.q)])

    def construct_right_indices(self):
        return np.nonzero(self.c[0] == 1)[0]

    def construct_left_indices(self):
        return np.nonzero(self.c[0] == -1)[0]

    def construct_main_indices(self):
        ct = self.c.T.astype(np.int64)
        return np.nonzero(np.abs(ct[:, 0]) + np.abs(ct[:, 1]) == 1)[0]

    def construct_lattice_velocity(self):
        return self.c.copy()

    def construct_lattice_weight(self):
        # lattice.py:147-190: 1/9 on the axes, 1/36 on the diagonals, 4/9 at rest (D2Q9)
        ct = self.c.T.astype(np.float64)
        w = np.full(self.q, 1.0 / 36.0)
        w[np.linalg.norm(ct, axis=1) < 1.1] = 1.0 / 9.0
        if self.name == "D2Q9":
            w[0] = 4.0 / 9.0
        return w

    def construct_lattice_moment(self):
        ct = self.c.T.astype(np.float64)
        cc = np.zeros((self.q, 3))
        k = 0
        for a in range(2):
            for b in range(a, 2):
                cc[:, k] = ct[:, a] * ct[:, b]
                k += 1
        return cc

    def __str__(self):
        return self.name


class LatticeD2Q8(Lattice):
    def __init__(self, precision="f32/f32"):
        super().__init__("D2Q8", precision)
        self.i_s = np.arange(8, dtype=np.int8)
        self.im = 2
        self.ik = 3


class LatticeD2Q9(Lattice):
    def __init__(self, precision="f32/f32"):
        super().__init__("D2Q9", precision)
        self.cs = np.sqrt(3) / 3.0
        self.cs2 = 1.0 / 3.0
        self.inv_cs2 = 3.0
        self.i_s = np.arange(9, dtype=np.int8)
        self.im = 3
        self.ik = 3
