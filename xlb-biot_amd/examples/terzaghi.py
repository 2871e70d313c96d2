"""Terzaghi's consolidation (quasi-2-D column): the study of
examples/poroelasticity/Terzaghi.py on this package's solvers.

Same command line (``-a -n -e -l -r -c -g --plots/--no-plots -s -v``), parameters
(E = 1, nu = 0.8, c0 = 1, L = 1 through ``Parameters``; beta = 1/4), corner-fixed
walls (left: fixed and impermeable, right: loaded and drained), series solutions
of the subsidence and the pressure (101 terms), result folder
``./results/poroelasticity/Terzaghi`` and ``.npz`` error histories (file name and
keys of Terzaghi.py:308-348).  ``-l -1`` is the paper's multigrid sweep
(runTerzaghi.sh): n x n cells, log2(n) levels, ``-e`` V-cycles per step.

    python xlb-biot_amd/examples/terzaghi.py -n 4 -r 5 -l -1 -a 1.0 -c centered --no-plots -e 1
"""
import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from common import Figures, depth_and_steps, print_convergence, refinement_list, tag
from src.darcy import ADR_AntiBounceBack, ADR_Neumann_BounceBack, AdvectionDiffusionReactionBGK
from src.elasticity import ElasticityAntiBounceBack, ElasticityBounceBackFullway, LBMElasticity
from src.lattice import LatticeD2Q8, LatticeD2Q9
from src.poroelasticity import LBMPoroelasticity
from src.utils import Parameters, save_fields_vtk

OUTPUT_DIR = os.path.join(".", "results", "poroelasticity", "Terzaghi")

# Directions each BC drops from its boundary mask at the lower / upper corner cell of
# the left / right wall (Terzaghi.py:22-111): keys (right wall, upper corner).
ADR_CORNERS = {(False, False): (1, 4), (False, True): (2, 8), (True, False): (1, 7), (True, True): (2, 5)}
EL_CORNERS = {(False, False): (1, 5), (False, True): (3, 6), (True, False): (1, 4), (True, True): (3, 7)}


def corner_fixed(Base, corners):
    """``Base`` whose missing directions pair every direction with its opposite, with
    ``corners`` removed at the wall corners and everything outside the boundary mask
    cleared (the drivers' ``get_missing_indices`` overrides)."""

    class CornerFixed(Base):
        def get_missing_indices(self, boundaryMask):
            nbd = len(self.indices[0])
            imissing = np.tile(np.arange(self.lattice.q, dtype="uint8"), (nbd, 1))
            iknown = np.tile(np.asarray(self.lattice.opp_indices), (nbd, 1))
            xs, ys = np.asarray(self.indices[0]), np.asarray(self.indices[1])
            for (right, top), dirs in corners.items():
                at = (xs == (self.nx - 2 if right else 1)) & (ys == (self.ny - 1 if top else 0))
                for i in np.nonzero(at)[0]:
                    boundaryMask[i, list(dirs)] = False
            imissing[~boundaryMask] = 0
            iknown[~boundaryMask] = 0
            return imissing, iknown

    CornerFixed.__name__ = "CornerFixed" + Base.__name__
    return CornerFixed


ADR_BB = corner_fixed(ADR_Neumann_BounceBack, ADR_CORNERS)
ADR_AntiBB = corner_fixed(ADR_AntiBounceBack, ADR_CORNERS)
ElasticityAntiBB = corner_fixed(ElasticityAntiBounceBack, EL_CORNERS)


class TerzaghiDarcy(AdvectionDiffusionReactionBGK):
    def initialize_macroscopic_fields(self):
        return np.zeros((self.nx, self.ny, 1)), np.zeros((self.nx, self.ny, 2))

    def set_boundary_conditions(self):
        bb = self.boundingBoxIndices
        self.BCs.append(ADR_BB(tuple(bb["left"].T), self.gridInfo, self.precisionPolicy))        # no flow
        self.BCs.append(ADR_AntiBB(tuple(bb["right"].T), self.gridInfo, self.precisionPolicy))   # p = 0


class TerzaghiElasticity(LBMElasticity):
    def __init__(self, **kwargs):
        self._load = -kwargs["parameters"].get_nondimensional("force")
        super().__init__(**kwargs)

    def initialize_macroscopic_fields(self):
        return np.zeros((self.nx, self.ny, 2)), np.zeros((self.nx, self.ny, 3))

    def set_boundary_conditions(self):
        bb = self.boundingBoxIndices
        self.BCs.append(ElasticityBounceBackFullway(tuple(bb["left"].T), self.gridInfo, self.precisionPolicy))
        traction = np.stack((np.full((self.ny,), self._load), np.zeros((self.ny,))), axis=-1)
        self.BCs.append(ElasticityAntiBB(tuple(bb["right"].T), self.gridInfo, self.precisionPolicy, traction))
        for level in range(1, self.levels):
            gi = self.grid_info(level)
            bbl = self.bounding_box_indices(level)
            self.BCs_MG[level] = [ElasticityBounceBackFullway(tuple(bbl["left"].T), gi, self.precisionPolicy),
                                  ElasticityAntiBB(tuple(bbl["right"].T), gi, self.precisionPolicy)]


class TerzaghiPoroelasticity(LBMPoroelasticity):
    """Series solutions (Terzaghi.py:166-183) and the driver's error histories
    (Terzaghi.py:185-214)."""

    def __init__(self, **kwargs):
        super().__init__(TerzaghiElasticity, TerzaghiDarcy, **kwargs)
        prm = kwargs["parameters"]
        self._output_dir = kwargs.get("output_dir")
        self._save_vtk = kwargs.get("save_vtk", False)
        self._j = np.arange(kwargs.get("N_exact", 100) + 1)
        self._deta, self._dx, self._dt, self._dp = (prm.get_unit(u) for u in ("H", "L", "T", "P"))
        self._eta_inf = prm.get_nondimensional("etaInf")

    def pressure(self, timestep):
        nx, ny = self.darcy.nx, self.darcy.ny
        if timestep <= 0:
            return np.ones((nx, ny, 1))
        j = self._j
        x = np.arange(nx)[:, None]
        terms = ((-1.0) ** j / (2 * j + 1)) * np.cos((2 * j + 1) * 0.5 * np.pi * self._dx * (x - 0.5)) \
            * np.exp((2 * j + 1) ** 2 * (-0.25 * np.pi ** 2 * timestep * self._dt))
        col = 4.0 / np.pi * terms.sum(axis=1)
        return np.broadcast_to(col[:, None, None], (nx, ny, 1)).copy()

    def subsidence(self, timestep):
        if timestep <= 0:
            return 1.0
        j = self._j
        s = np.sum(np.exp((2 * j + 1) ** 2 * (-0.25 * np.pi ** 2 * timestep * self._dt)) / (2 * j + 1) ** 2)
        return self._eta_inf + (1.0 - self._eta_inf) * 8.0 / np.pi ** 2 * s

    def run(self, t_max):
        n = t_max // self.ioRate + 1
        self._l2_subsidence, self._err_l2_sub = np.empty(n), np.empty(n)
        self._l2_pressure, self._err_l2_pres = np.empty(n), np.empty(n)
        return super().run(t_max)

    def output_data(self, **kwargs):
        t = kwargs["timestep"]
        i = t // self.ioRate
        eta = np.array(kwargs["eta"])[0]
        pressure = np.array(kwargs["pressure"])[0]
        sub = self.subsidence(t)
        self._l2_subsidence[i] = sub
        self._err_l2_sub[i] = np.mean(eta[-2, :, 0]) + sub  # the subsidence has the opposite sign
        p_ex = self.pressure(t)
        self._l2_pressure[i] = np.linalg.norm(p_ex[1:-1])
        self._err_l2_pres[i] = np.linalg.norm((pressure - p_ex)[1:-1])
        if self._save_vtk:
            sigma = np.array(kwargs["sigma"])[0]
            fields = {"eta_x": eta[..., 0] * self._deta, "eta_y": eta[..., 1] * self._deta,
                      "sigma_xx": sigma[..., 0] * self._dp / self.epsilon,
                      "sigma_xy": sigma[..., 1] * self._dp / self.epsilon,
                      "sigma_yy": sigma[..., 2] * self._dp / self.epsilon, "p": pressure[..., 0] * self._dp,
                      "p_ex": p_ex[..., 0] * self._dp, "diff_p": (pressure - p_ex)[..., 0] * self._dp}
            save_fields_vtk(i, fields, self._output_dir)

    def error_data(self):
        """The driver's ``.npz`` record (Terzaghi.py:336-340)."""
        return {"err_l2_sub": self._err_l2_sub, "err_l2_pres": self._err_l2_pres,
                "rel_err_sub": np.sqrt(np.sum(self._err_l2_sub ** 2) / np.sum(self._l2_subsidence ** 2)),
                "rel_err_pres": np.sqrt(np.sum(self._err_l2_pres ** 2) / np.sum(self._l2_pressure ** 2))}


def parameters(alpha):
    """The driver's dimensional set-up and its derived quantities (Terzaghi.py:247-277)."""
    p = Parameters()
    p.add_units(L=1.0, H=1.0, T=1.0, P=1.0)
    p.add_parameters(L_tot=(1e0, {"L": 1}), E=(1.0, {"P": 1}), nu=(0.8, {}), c0=(1e0, {"P": -1}), alpha=(alpha, {}))
    E, nu = p.get("E"), p.get("nu")
    p.set_parameter("lambda", E * nu / (1 - nu ** 2), {"P": 1})
    p.set_parameter("mu", 0.5 * E / (1 + nu), {"P": 1})
    a, c0, lam, mu = p.get("alpha"), p.get("c0"), p.get("lambda"), p.get("mu")
    p.add_parameters(force=(a / c0 + (lam + 2 * mu) / a, {"L": -1, "H": 1, "P": 1}),
                     Kpm=(lam + 2 * mu + a ** 2 / c0, {"P": 1}))
    force, Kpm, L = p.get("force"), p.get("Kpm"), p.get("L_tot")
    p.add_parameters(p0=(a * force / c0 / Kpm, {"P": 1}), eta0=(force * L / Kpm, {"H": 1}),
                     etaInf=(force * L / (lam + 2 * mu), {"H": 1}),
                     kappa=(Kpm / (lam + 2 * mu), {"L": 2, "T": -1, "P": -1}))
    p.set_parameter("c_f", p.get("kappa") * (lam + 2 * mu) / c0 / Kpm, {"L": 2, "T": -1})
    p.set_parameter("T_tot", L ** 2 / p.get("c_f"), {"T": 1})
    return p


def npz_name(output_dir, alpha, nx, ne, levels, coupling, gradient):
    return os.path.join(output_dir, f"error_a{alpha:.3f}_nx{nx:03d}_ne{ne:03d}_l{levels:03d}_{tag(coupling, gradient)}.npz")


def study(nx=20, refinements=0, levels=1, elastic_steps=1.0, coupling="centered", gradient=False, alpha=1.0,
          plots=False, save_fig=False, vtk=False, output_dir=OUTPUT_DIR, print_info_rate=1000, verbose=False):
    """One refinement study; returns ``{nx: record}`` (the ``.npz`` contents)."""
    precision = "f64/f64"
    prm = parameters(alpha)
    if verbose:
        print("### DIMENSIONAL " + "#" * 44)
        prm.print(nondimensional=False)
    prm.add_units(L=prm.get("L_tot"), H=prm.get("eta0"), T=prm.get("T_tot"), P=prm.get("p0"))
    for k, v in {"L_tot": 1.0, "c0": 1.0, "p0": 1.0, "eta0": 1.0}.items():
        if abs(prm.get_nondimensional(k) - v) > 1e-6 * abs(v):
            raise RuntimeError(f"Wrong conversion of parameter '{k}'.")
    output_dir = os.path.abspath(output_dir)
    os.makedirs(output_dir, exist_ok=True)
    fig = Figures(plots)
    nx_list = refinement_list(nx, refinements, levels)
    results = {}
    for n in nx_list:
        beta = 0.25
        prm.add_units(L=prm.get("L_tot") / n, H=prm.get("eta0"), T=prm.get("T_tot") * beta / n ** 2, P=prm.get("p0"))
        if verbose:
            prm.print()
        nt = int(prm.get_nondimensional("T_tot"))
        depth, ne = depth_and_steps(n, levels, elastic_steps)
        ny = n if levels < 1 else 4
        io_rate = max(1, nt // 400) if vtk else 1
        fname = npz_name(output_dir, alpha, n, ne, depth, coupling, gradient)
        data = None
        if not vtk:
            try:
                with np.load(fname) as z:
                    data = dict(z)
            except (OSError, ValueError):
                data = None
        if data is None:
            sim = TerzaghiPoroelasticity(**{
                "lattice_elasticity": LatticeD2Q8(precision), "lattice_darcy": LatticeD2Q9(precision),
                "elasticity_steps": ne, "elasticity_MG_levels": depth, "MG_BCs": {"left": True, "right": True},
                "MG_n_smoothing": 4, "coupling_method": coupling, "alpha": prm.get_nondimensional("alpha"),
                "lambda": prm.get_nondimensional("lambda"), "mu": prm.get_nondimensional("mu"),
                "epsilon": prm.get_unit("L") / prm.get_unit("H"),
                "omega": 1.0 / (3 * prm.get_nondimensional("kappa") + 0.5), "parameters": prm,
                "nx": n + 2, "ny": ny, "nz": 0, "precision": precision, "io_rate": io_rate, "save_vtk": vtk,
                "print_info_rate": print_info_rate, "output_dir": output_dir})
            sim.run(nt)
            data = sim.error_data()
            if not vtk:
                np.savez_compressed(fname, **data)
        results[n] = data
        if fig.plt is not None:
            times = np.arange(0.0, nt + 1, io_rate)
            fig.plt.plot(times, data["err_l2_sub"], color="#3070B3", label="L2 subs.")
            fig.plt.plot(times, data["err_l2_pres"] / n, color="#EF9067", label="e^k_p")
            fig.plt.yscale("log")
            fig.plt.xlabel("Time Step")
            fig.plt.ylabel("Error")
            fig.plt.legend(loc="upper right")
            fig.plt.title(f"alpha = {alpha}, N_x = {n}, N_E = {ne}, {coupling}")
            fig.finish(os.path.join(output_dir, f"a{alpha:.3f}_nx{n:03d}_ne{ne:03d}_l{depth:03d}_"
                                                f"{tag(coupling, gradient)}") if save_fig else None)
    if len(nx_list) > 1:
        names = ["rel_err_sub", "rel_err_pres"]
        print_convergence(names, nx_list, {k: [results[n][k] for n in nx_list] for k in names}, 1)
    return results


def main(argv=None):
    ap = argparse.ArgumentParser(description="Quasi-2D Terzaghi poroelastic simulation using LBM (MI355X).")
    ap.add_argument("-a", "--alpha", type=float, default=1.0, help="Biot-Willis coefficient (default: 1)")
    ap.add_argument("-n", "--nx", type=int, default=20, help="cells in x (default: 20)")
    ap.add_argument("-e", "--elastic_steps", type=float, default=1.0,
                    help="relative number of elasticity steps per time step (default: 1.0)")
    ap.add_argument("-l", "--levels", type=int, default=1, help="multigrid levels for elasticity (default: 1)")
    ap.add_argument("-r", "--refinements", type=int, default=0, help="refinements of the study (default: 0)")
    ap.add_argument("-c", "--coupling", type=str, default="centered", choices=["explicit", "implicit", "centered"])
    ap.add_argument("-g", "--gradient", action=argparse.BooleanOptionalAction, default=False,
                    help="marks the result files as alternative-gradient runs (default: False)")
    ap.add_argument("-p", "--plots", action=argparse.BooleanOptionalAction, default=True)
    ap.add_argument("-s", "--save_fig", action=argparse.BooleanOptionalAction, default=False)
    ap.add_argument("-v", "--vtk", action=argparse.BooleanOptionalAction, default=False)
    a = ap.parse_args(argv)
    study(a.nx, a.refinements, a.levels, a.elastic_steps, a.coupling, a.gradient, a.alpha, a.plots, a.save_fig,
          a.vtk, verbose=True)


if __name__ == "__main__":
    main()
