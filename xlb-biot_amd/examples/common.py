"""Shared pieces of the study drivers: cell-centred mode sums, the refinement and
depth rules of the drivers' command lines, and plotting that degrades to a
printed table when no figure is wanted."""
import os
import sys

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if PKG not in sys.path:
    sys.path.insert(0, PKG)


def mode_sum(nx, ny, kls, scale=lambda k, l: 1.0):
    """``(nx, ny, 1)`` sum of separable modes at the cell centres (the drivers'
    ``sumFunction``, examples/poroelasticity/periodic.py:23-31): ``kls[j]`` lists
    wave-number pairs ``k1, l1, k2, l2, ...`` of the products cos cos, cos sin,
    sin cos, sin sin (j = 0..3), each term ``scale(k, l) f(k pi x') g(l pi y')`` with
    ``x' = (x + 1/2) / nx``."""
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    pairs = [(np.cos, np.cos), (np.cos, np.sin), (np.sin, np.cos), (np.sin, np.sin)]
    out = np.zeros((nx, ny))
    for (fx, fy), ks in zip(pairs, kls):
        for k, l in zip(ks[0::2], ks[1::2]):
            out += scale(k, l) * np.outer(fx(k * np.pi * x), fy(l * np.pi * y))
    return out[..., None]


def refinement_list(nx, refinements, levels):
    """Grid sizes of a study: doubling with automatic multigrid depth (``levels`` < 0),
    else ``nx (r + 1)`` (periodic.py:192-195, Terzaghi.py:287-290)."""
    if levels < 0:
        return [nx * 2 ** r for r in range(refinements + 1)]
    return [nx * (r + 1) for r in range(refinements + 1)]


def depth_and_steps(nx, levels, elastic_steps):
    """The drivers' multigrid depth and elasticity steps per physical step: with
    ``levels`` < 1 the deepest hierarchy ``nx`` allows and ``int(elastic_steps)``
    V-cycles, else ``levels`` and ``max(1, round(nx elastic_steps))`` steps
    (periodic.py:209-216, Terzaghi.py:298-306)."""
    if levels < 1:
        depth = int(np.log2(nx))
        while nx % 2 ** (depth - 1) != 0:
            depth -= 1
        return depth, int(elastic_steps)
    return levels, max(1, round(nx * elastic_steps))


def tag(coupling, gradient):
    return f"{coupling}{'_altgrad' if gradient else ''}"


class Figures:
    """Error-history and convergence figures (matplotlib, off-screen); ``save`` writes
    ``<stem>.eps`` (and ``<stem>.tex`` when tikzplotlib is installed), ``show``
    leaves nothing on a headless host."""

    def __init__(self, enabled):
        self.enabled = enabled
        self.plt = None
        if enabled:
            import matplotlib
            if not os.environ.get("DISPLAY"):
                matplotlib.use("Agg")
            from matplotlib import pyplot
            self.plt = pyplot

    def finish(self, stem=None):
        if self.plt is None:
            return
        if stem is not None:
            self.plt.savefig(stem + ".eps")
            try:
                from tikzplotlib import save as save_tikz
                save_tikz(stem + ".tex")
            except ImportError:
                pass
            self.plt.close()
        else:
            self.plt.show()


def print_convergence(names, nx_list, table, order_ref):
    """The refinement table with observed orders between successive grids."""
    print(f"{'nx':>6} " + " ".join(f"{n:>14}" for n in names))
    for i, nx in enumerate(nx_list):
        row = f"{nx:>6} " + " ".join(f"{table[n][i]:>14.6e}" for n in names)
        if i:
            rates = [np.log(table[n][i - 1] / table[n][i]) / np.log(nx_list[i] / nx_list[i - 1]) for n in names]
            row += "   orders " + " ".join(f"{r:5.2f}" for r in rates)
        print(row)
    print(f"(reference slope O(dx^{order_ref}))")
