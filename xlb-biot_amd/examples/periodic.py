"""Periodic poroelasticity with a manufactured time-dependent solution: the study of
examples/poroelasticity/periodic.py on this package's solvers.

Same command line (``-a -n -e -l -r -c -g --plots/--no-plots -s -v``), material
(E = 0.11, nu = 0.8, K = 0.1), step counts (nx^2 physical steps, io every step),
result folder ``./results/poroelasticity/periodic`` and ``.npz`` error histories
(file name and keys of periodic.py:217-259; an existing file is read instead of
running again).  The forcing, source and exact fields follow periodic.py:48-149.

    python xlb-biot_amd/examples/periodic.py -n 10 -r 4 -a 0.5 -c centered --no-plots -e 1.0
"""
import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from common import Figures, depth_and_steps, mode_sum, print_convergence, refinement_list, tag
from src.darcy import AdvectionDiffusionReactionBGK
from src.elasticity import LBMElasticity
from src.lattice import LatticeD2Q8, LatticeD2Q9
from src.poroelasticity import LBMPoroelasticity
from src.utils import save_fields_vtk

E_MOD, NU, K_PERM = 0.11, 0.8, 0.1
OUTPUT_DIR = os.path.join(".", "results", "poroelasticity", "periodic")
HISTORY_KEYS = ("eta", "sigma", "p", "p_grad", "dtdiv_eta")


class PeriodicDarcy(AdvectionDiffusionReactionBGK):
    def initialize_macroscopic_fields(self):
        return np.zeros((self.nx, self.ny, 1)), np.zeros((self.nx, self.ny, 2))


class PeriodicElasticity(LBMElasticity):
    def initialize_macroscopic_fields(self):
        return np.zeros((self.nx, self.ny, 2)), np.zeros((self.nx, self.ny, 3))


class PeriodicPoroelasticity(LBMPoroelasticity):
    """Body force and Darcy source of the manufactured solution (periodic.py:48-63)
    and the driver's error histories (periodic.py:65-149)."""

    def __init__(self, **kwargs):
        # spatial patterns of every field (the time dependence is a scalar factor):
        # cos(2 pi x) sin(2 pi y), sin cos, sin sin, cos cos; the hooks run during the
        # base construction already
        nx, ny = kwargs["nx"], kwargs["ny"]
        self._cs = mode_sum(nx, ny, [[], [2, 2], [], []])
        self._sc = mode_sum(nx, ny, [[], [], [2, 2], []])
        self._ss = mode_sum(nx, ny, [[], [], [], [2, 2]])
        self._cc = mode_sum(nx, ny, [[2, 2], [], [], []])
        super().__init__(PeriodicElasticity, PeriodicDarcy, **kwargs)
        self._output_dir = kwargs.get("output_dir")
        self._save_vtk = kwargs.get("save_vtk", False)

    def _k(self):
        return (1.0 / float(np.mean(np.asarray(self.darcy.omega))) - 0.5) / 3

    def get_elastic_force(self, timestep):
        tf = 1 - np.exp(-8 * np.pi ** 2 * self._k() * timestep / self.elasticity.nx ** 2)
        fx = self._cs * (self.epsilon ** 2 * tf * np.pi ** 2 * 8 * self.elasticity.mu)
        return np.concatenate((fx, np.zeros_like(fx)), axis=-1)

    def get_darcy_source(self, timestep):
        K = self._k()
        tf = np.exp(-8 * np.pi ** 2 * K * timestep / self.elasticity.nx ** 2)
        factor = -np.pi ** 3 * K / self.darcy.nx ** 2 * (
            128 * self.alpha * tf + (128 * self.elasticity.lamda + 240 * self.elasticity.mu) / self.alpha)
        return self._ss * factor

    def run(self, t_max):
        n = t_max // self.ioRate + 1
        self._hist = {f"{kind}_{key}": np.empty(n) for key in HISTORY_KEYS
                      for kind in ("err_l2", "err_lInf", "l2_exact", "lInf_exact")}
        return super().run(t_max)

    def _record(self, i, key, diff, exact):
        h = self._hist
        h[f"l2_exact_{key}"][i] = np.linalg.norm(exact)
        h[f"lInf_exact_{key}"][i] = np.abs(exact).max()
        h[f"err_l2_{key}"][i] = np.linalg.norm(diff)
        h[f"err_lInf_{key}"][i] = np.abs(diff).max()

    def output_data(self, **kwargs):
        K = self._k()
        lam, mu, a = self.elasticity.lamda, self.elasticity.mu, self.alpha
        t = kwargs["timestep"]
        i = t // self.ioRate
        eta = np.array(kwargs["eta"])[0]
        dt_div_eta = np.array(kwargs["dt_div_eta"])[0] / self.epsilon ** 3
        sigma = np.array(kwargs["sigma"])[0] / self.epsilon
        pressure = np.array(kwargs["pressure"])[0]
        pressure_grad = np.array(kwargs["pressure_grad"])[0] / self.epsilon
        tf = 1 - np.exp(-8 * np.pi ** 2 * K * t / self.elasticity.nx ** 2)
        eta_ex = np.concatenate((4.5 * tf * self._cs, 3.5 * tf * self._sc), axis=-1)
        dtdiv_ex = -128 * np.pi ** 3 * K * np.exp(-8 * np.pi ** 2 * K * t / self.elasticity.nx ** 2) * self._ss
        sigma_ex = np.concatenate((-np.pi * tf * (16 * lam + 18 * mu) * self._ss, np.pi * tf * 16 * mu * self._cc,
                                   -np.pi * tf * (16 * lam + 14 * mu) * self._ss), axis=-1)
        p_ex = -np.pi / a * tf * (16 * lam + 30 * mu) * self._ss
        g = -2 * np.pi ** 2 / a * tf * (16 * lam + 30 * mu)
        p_grad_ex = np.concatenate((g * self._cs, g * self._sc), axis=-1)
        self._record(i, "eta", eta - eta_ex, eta_ex)
        self._record(i, "dtdiv_eta", dt_div_eta - dtdiv_ex, dtdiv_ex)
        self._record(i, "sigma", sigma - sigma_ex, sigma_ex)
        self._record(i, "p", pressure - p_ex, p_ex)
        self._record(i, "p_grad", pressure_grad - p_grad_ex, p_grad_ex)
        if self._save_vtk:
            fields = {"eta_x": eta[..., 0], "eta_y": eta[..., 1], "sigma_xx": sigma[..., 0],
                      "sigma_xy": sigma[..., 1], "sigma_yy": sigma[..., 2], "p": pressure[..., 0],
                      "p_ex": p_ex[..., 0], "diff_p": (pressure - p_ex)[..., 0], "dp_dx": pressure_grad[..., 0],
                      "dp_dy": pressure_grad[..., 1]}
            save_fields_vtk(t, fields, self._output_dir)

    def error_data(self):
        """The driver's ``.npz`` record (periodic.py:234-246)."""
        h = self._hist
        data = {}
        for key in HISTORY_KEYS:
            data[f"err_l2_{key}"] = h[f"err_l2_{key}"]
            data[f"err_lInf_{key}"] = h[f"err_lInf_{key}"]
        for key in HISTORY_KEYS:
            data[f"rel_err_l2_{key}"] = np.sqrt(np.sum(h[f"err_l2_{key}"] ** 2) / np.sum(h[f"l2_exact_{key}"] ** 2))
            data[f"rel_err_lInf_{key}"] = np.max(h[f"err_lInf_{key}"]) / np.max(h[f"lInf_exact_{key}"])
        return data


def npz_name(output_dir, alpha, nx, ne, levels, coupling, gradient):
    return os.path.join(output_dir, f"error_E{E_MOD:.3f}_nu{NU:.3f}_K{K_PERM:.3f}_a{alpha:.3f}_nx{nx:03d}_ne{ne:03d}"
                                    f"_l{levels:03d}_{tag(coupling, gradient)}.npz")


def study(nx=20, refinements=0, levels=1, elastic_steps=1.0, coupling="centered", gradient=False, alpha=1.0,
          plots=False, save_fig=False, vtk=False, output_dir=OUTPUT_DIR, print_info_rate=1000):
    """One refinement study; returns ``{nx: record}`` (the ``.npz`` contents)."""
    precision = "f64/f64"
    lam = E_MOD * NU / (1 - NU ** 2)
    mu = 0.5 * E_MOD / (1 + NU)
    omega = 1.0 / (3 * K_PERM + 0.5)
    output_dir = os.path.abspath(output_dir)
    os.makedirs(output_dir, exist_ok=True)
    fig = Figures(plots)
    nx_list = refinement_list(nx, refinements, levels)
    results = {}
    for n in nx_list:
        depth, ne = depth_and_steps(n, levels, elastic_steps)
        nt = n ** 2
        io_rate = max(1, nt // 100) if vtk else 1
        fname = npz_name(output_dir, alpha, n, ne, depth, coupling, gradient)
        try:
            with np.load(fname) as z:
                data = dict(z)
        except (OSError, ValueError):
            sim = PeriodicPoroelasticity(**{
                "lattice_elasticity": LatticeD2Q8(precision), "lattice_darcy": LatticeD2Q9(precision),
                "elasticity_steps": ne, "elasticity_MG_levels": depth, "coupling_method": coupling,
                "alternative_gradient": gradient, "alpha": alpha, "lambda": lam, "mu": mu, "omega": omega,
                "nx": n, "ny": n, "nz": 0, "precision": precision, "io_rate": io_rate, "save_vtk": vtk,
                "print_info_rate": print_info_rate, "output_dir": output_dir})
            sim.run(nt)
            data = sim.error_data()
            np.savez_compressed(fname, **data)
        results[n] = data
        if fig.plt is not None:
            times = np.arange(0.0, nt + 1, io_rate)
            for key, col in (("eta", "#3070B3"), ("sigma", "#EF9067"), ("p", "#A2AD00")):
                fig.plt.plot(times, data[f"err_l2_{key}"] / n, "-", color=col, label=f"e^k_{key}")
                fig.plt.plot(times, data[f"err_lInf_{key}"], "--", color=col, label=f"Linf {key}")
            fig.plt.yscale("log")
            fig.plt.xlabel("Time Step k")
            fig.plt.ylabel("Error")
            fig.plt.legend(loc="upper right")
            fig.plt.title(f"alpha = {alpha}, N_x = {n}, N_E = {ne}, {coupling}")
            fig.finish(os.path.join(output_dir, f"a{alpha:.3f}_nx{n:03d}_ne{ne:03d}_l{depth:03d}_"
                                                f"{tag(coupling, gradient)}") if save_fig else None)
    if len(nx_list) > 1:
        names = ["rel_err_l2_eta", "rel_err_l2_sigma", "rel_err_l2_p"]
        print_convergence(names, nx_list, {k: [results[n][k] for n in nx_list] for k in names}, 2)
    return results


def main(argv=None):
    ap = argparse.ArgumentParser(description="2D periodic poroelastic simulation using LBM (MI355X).")
    ap.add_argument("-a", "--alpha", type=float, default=1.0, help="Biot-Willis coefficient (default: 1.0)")
    ap.add_argument("-n", "--nx", type=int, default=20, help="cells in x (default: 20)")
    ap.add_argument("-e", "--elastic_steps", type=float, default=1.0,
                    help="relative number of elasticity steps per time step (default: 1.0)")
    ap.add_argument("-l", "--levels", type=int, default=1, help="multigrid levels for elasticity (default: 1)")
    ap.add_argument("-r", "--refinements", type=int, default=0, help="refinements of the study (default: 0)")
    ap.add_argument("-c", "--coupling", type=str, default="centered", choices=["explicit", "implicit", "centered"])
    ap.add_argument("-g", "--gradient", action=argparse.BooleanOptionalAction, default=False,
                    help="use the alternative pressure gradient (default: False)")
    ap.add_argument("-p", "--plots", action=argparse.BooleanOptionalAction, default=True)
    ap.add_argument("-s", "--save_fig", action=argparse.BooleanOptionalAction, default=False)
    ap.add_argument("-v", "--vtk", action=argparse.BooleanOptionalAction, default=False)
    a = ap.parse_args(argv)
    study(a.nx, a.refinements, a.levels, a.elastic_steps, a.coupling, a.gradient, a.alpha, a.plots, a.save_fig,
          a.vtk)


if __name__ == "__main__":
    main()
