"""Hand-derived wall tables of the elastic anti-bounce-back (traction, reference
elasticity.py:1648-1750) and the ADR anti-bounce-back (Dirichlet, darcy.py:298-392)
conditions, checked against the oracle's literal scatter passes on 4 x 6 grids
with one wall at a time.

The tables follow from the reference's definitions by hand:

* a halfway condition acts on the first fluid layer (the wall layer shifted by
  its normal); direction d is missing there when its upstream cell x - c_d is the
  wall, i.e. c_d points away from the wall;
* ElasticityAntiBounceBack writes ``f[missingMask] = -fin[knownMask]`` with
  ``knownMask = missingMask[:, opp]``: boolean masks pair in row-major (increasing
  direction) order, not by opposites, and the traction ``T`` is added per normal
  (elasticity.py:1710-1727);
* ADR_AntiBounceBack writes ``f[imissing] = -fin[iknown] + 2 w rho`` with iknown
  the opposites; the non-missing entries of both index arrays are 0, so the rest
  population is set to ``-fin[0]`` and receives ``2 w_0 rho`` once per
  non-missing direction (``.at[].add`` accumulates duplicate indices).

Only wall cells away from the periodic faces are checked (next to them the halo,
marked solid, adds missing diagonals)."""
import numpy as np
import pytest
import torch

import problems as P
from oracle import bc as obc

NX, NY = 4, 6
FIRST = {"left": ("x", 1), "right": ("x", NX - 2), "bottom": ("y", 1), "top": ("y", NY - 2)}

# D2Q8, c = (1,0) (0,1) (-1,0) (0,-1) (1,1) (-1,1) (-1,-1) (1,-1), opp = 2 3 0 1 6 7 4 5.
# missing direction <- known direction (row-major pairing of the two masks)
EABB_PAIRS = {"left": {0: 2, 4: 5, 7: 6}, "right": {2: 0, 5: 4, 6: 7},
              "bottom": {1: 3, 4: 6, 5: 7}, "top": {3: 1, 6: 4, 7: 5}}
# traction added: direction -> (component of T, factor)
EABB_T = {"left": {0: (0, 1.0), 4: (1, 0.5), 7: (1, -0.5)}, "right": {2: (0, -1.0), 5: (1, 0.5), 6: (1, -0.5)},
          "bottom": {1: (1, 1.0), 4: (0, 0.5), 5: (0, -0.5)}, "top": {3: (1, -1.0), 6: (0, -0.5), 7: (0, 0.5)}}

# D2Q9, cx = 0 0 0 1 -1 1 -1 1 -1, cy = 0 1 -1 0 1 -1 0 1 -1, opp = 0 2 1 6 5 4 3 8 7,
# w = 4/9, 1/9, 1/9, 1/9, 1/36, 1/36, 1/9, 1/36, 1/36
ADR_PAIRS = {"left": {3: 6, 5: 4, 7: 8}, "right": {4: 5, 6: 3, 8: 7},
             "bottom": {1: 2, 4: 5, 7: 8}, "top": {2: 1, 5: 4, 8: 7}}
W9 = [4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 9, 1 / 36, 1 / 36]


def _wall_cells(wall):
    axis, k = FIRST[wall]
    if axis == "x":
        return [(k, y) for y in range(1, NY - 1)]
    return [(x, k) for x in range(1, NX - 1)]


def _oracle_elasticity(wall, T):
    m = P.modules("oracle")

    class One(m["Elasticity"]):
        def set_boundary_conditions(self):
            self.BCs.append(m["EABB"](tuple(self.boundingBoxIndices[wall].T), self.gridInfo, self.precisionPolicy,
                                      None if T is None else np.asarray(T)))

    return One(**{"lambda": 0.7, "mu": 0.3, "nx": NX, "ny": NY, "nz": 0, "precision": "f64/f64",
                  "lattice": m["D2Q8"]()})


def _oracle_darcy(wall, rho):
    m = P.modules("oracle")

    class One(m["Darcy"]):
        def set_boundary_conditions(self):
            self.BCs.append(m["ADR_ABB"](tuple(self.boundingBoxIndices[wall].T), self.gridInfo, self.precisionPolicy,
                                         rho))

    return One(**{"omega": 1.2, "nx": NX, "ny": NY, "nz": 0, "precision": "f64/f64", "lattice": m["D2Q9"]()})


@pytest.mark.parametrize("T", [None, (3e-3, -7e-3)])
@pytest.mark.parametrize("wall", ["left", "right", "bottom", "top"])
def test_elastic_antibounceback_table(wall, T):
    el = _oracle_elasticity(wall, T)
    rng = np.random.default_rng(11)
    fout = rng.standard_normal((NX, NY, 8))
    fin = rng.standard_normal((NX, NY, 8))
    out = obc.apply_bc(el.BCs, torch.as_tensor(fout), torch.as_tensor(fin), "PostStreaming").numpy()
    touched = np.zeros((NX, NY), dtype=bool)
    for x, y in _wall_cells(wall):
        exp = fout[x, y].copy()
        for miss, known in EABB_PAIRS[wall].items():
            exp[miss] = -fin[x, y, known]
            if T is not None:
                comp, fac = EABB_T[wall][miss]
                exp[miss] += fac * T[comp]
        np.testing.assert_allclose(out[x, y], exp, rtol=0, atol=1e-15, err_msg=f"{wall} wall cell {(x, y)}")
        touched[x, y] = True
    # cells off the first fluid layer are left alone
    axis, k = FIRST[wall]
    for x in range(NX):
        for y in range(NY):
            if (axis == "x" and x != k) or (axis == "y" and y != k):
                np.testing.assert_array_equal(out[x, y], fout[x, y])


@pytest.mark.parametrize("rho", [None, 0.8])
@pytest.mark.parametrize("wall", ["left", "right", "bottom", "top"])
def test_adr_antibounceback_table(wall, rho):
    d = _oracle_darcy(wall, rho)
    rng = np.random.default_rng(12)
    fout = rng.standard_normal((NX, NY, 9))
    fin = rng.standard_normal((NX, NY, 9))
    out = obc.apply_bc(d.BCs, torch.as_tensor(fout), torch.as_tensor(fin), "PostStreaming").numpy()
    for x, y in _wall_cells(wall):
        exp = fout[x, y].copy()
        for miss, known in ADR_PAIRS[wall].items():
            exp[miss] = -fin[x, y, known] + (0.0 if rho is None else 2 * W9[miss] * rho)
        # the rest population: -fin[0], plus 2 w_0 rho for each of the 9 - 3 non-missing entries
        exp[0] = -fin[x, y, 0] + (0.0 if rho is None else 6 * 2 * W9[0] * rho)
        np.testing.assert_allclose(out[x, y], exp, rtol=0, atol=1e-15, err_msg=f"{wall} wall cell {(x, y)}")
    axis, k = FIRST[wall]
    for x in range(NX):
        for y in range(NY):
            if (axis == "x" and x != k) or (axis == "y" and y != k):
                np.testing.assert_array_equal(out[x, y], fout[x, y])


def test_row_major_pairing_mirrors_the_diagonals_of_the_side_walls():
    # on the left and right walls the row-major pairing takes each missing diagonal
    # from its mirror image across the wall, not from its opposite; on the bottom and
    # top walls the two coincide with the opposites
    opp8 = [2, 3, 0, 1, 6, 7, 4, 5]
    for wall, pairs in EABB_PAIRS.items():
        for m, k in pairs.items():
            if m < 4 or wall in ("bottom", "top"):
                assert k == opp8[m], (wall, m)
            else:
                assert k != opp8[m], (wall, m)
