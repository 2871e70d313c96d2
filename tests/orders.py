"""Convergence studies of the reference's example drivers, built on the GPU
solvers or on the oracle (``kind`` = "gpu" / "oracle").  Each study builds the
driver's set-up (manufactured or analytic solution, boundary conditions,
parameters, step counts, error norms) on a grid of ``n`` cells per side and
returns the driver's relative errors; :func:`order` fits the slope of the errors
against the grid spacing.  Which orders the drivers show:

* Darcy with anti-bounce-back (Dirichlet) walls, examples/darcy/dirichlet.py:
  errors over ~100 outputs of the decaying mode, plotted against O(nx^-1) and
  O(nx^-2);
* static elasticity with full-way (Dirichlet) or anti-bounce-back traction
  (Neumann) walls, examples/elasticity/{dirichlet,neumann}.py: errors of the
  final state after 3 n^2 steps, plotted against O(nx^-1) and O(nx^-2);
* periodic Biot with the manufactured time-dependent solution,
  examples/poroelasticity/periodic.py: errors accumulated over every physical
  step, plotted against O(dx^2);
* Terzaghi's consolidation, examples/poroelasticity/Terzaghi.py: subsidence and
  pressure errors over every physical step against the series solution, plotted
  against O(dx).
"""
import numpy as np

import problems as P


def _host(a):
    if hasattr(a, "detach"):
        a = a.detach().cpu()
    return np.asarray(a, dtype=np.float64)


def order(ns, errs):
    """Least-squares slope of log(error) against log(1/n)."""
    return float(np.polyfit(np.log(1.0 / np.asarray(ns, float)), np.log(np.asarray(errs, float)), 1)[0])


def _rel(err_sq, ref_sq):
    return float(np.sqrt(np.sum(err_sq) / np.sum(ref_sq)))


# ---------------------------------------------------------------------------
# Darcy, Dirichlet walls

def darcy_dirichlet(kind, n, precision="f64/f64"):
    """examples/darcy/dirichlet.py:22-79, 89-141: K = 0.1, the modes sin(2 pi x)
    sin(4 pi y) + sin(4 pi x) sin(2 pi y) decaying to half their amplitude under a
    static source, anti-bounce-back (p = 0) on all four walls, n^2 steps, output
    every n^2 / 100 steps; L2 and max errors relative to the exact solution,
    accumulated over the outputs."""
    m = P.modules(kind)
    ABB = m["ADR_ABB"]
    K = 0.1
    kls = [[], [], [], [2, 4, 4, 2]]
    out = {}

    class Dirichlet(m["Darcy"]):
        def get_source(self):
            Kd = (1.0 / self.omega - 0.5) / 3
            s = P.sum_function(self.nx, self.ny, kls,
                               lambda k, l: 0.5 * Kd * (k ** 2 + l ** 2) * np.pi ** 2 / (self.nx - 2) ** 2, walls=True)
            s[(0, -1), :] = 0
            s[:, (0, -1)] = 0
            return s

        def initialize_macroscopic_fields(self):
            p = P.sum_function(self.nx, self.ny, kls, walls=True) - 0.5 * self.get_source()
            return p, np.zeros((self.nx, self.ny, 2))

        def set_boundary_conditions(self):
            bb = self.boundingBoxIndices
            walls = np.concatenate((bb["left"], bb["right"], bb["top"], bb["bottom"]))
            self.BCs.append(ABB(tuple(walls.T), self.gridInfo, self.precisionPolicy))

        def output_data(self, **kw):
            p = np.array(kw["rho"])[0]
            t = kw["timestep"]
            Kt = (1.0 / self.omega - 0.5) / 3 * t / self.nx ** 2
            pe = P.sum_function(self.nx, self.ny, kls,
                                lambda k, l: 0.5 + 0.5 * np.exp(-Kt * (k ** 2 + l ** 2) * np.pi ** 2), walls=True)
            d, e = (p - pe)[1:-1, 1:-1].ravel(), pe[1:-1, 1:-1].ravel()
            out[t // self.ioRate] = (np.sum(d ** 2), np.sum(e ** 2), np.abs(d).max(), np.abs(e).max())

    nt = n * n
    sim = Dirichlet(lattice=m["D2Q9"](precision), omega=1.0 / (3 * K + 0.5), nx=n + 2, ny=n + 2, nz=0,
                    precision=precision, io_rate=max(1, nt // 100), print_info_rate=0)
    sim.run(nt)
    v = np.array([out[k] for k in sorted(out)])
    return {"l2": _rel(v[:, 0], v[:, 1]), "linf": float(v[:, 2].max() / v[:, 3].max())}


# ---------------------------------------------------------------------------
# static elasticity, Dirichlet or Neumann walls

E_EL, NU_EL = 0.11, 0.8   # the elasticity drivers' material (examples/elasticity/*.py)


def elasticity_walls(kind, n, walls="dirichlet", precision="f64/f64", steps_factor=3):
    """examples/elasticity/{dirichlet,neumann}.py: E = 0.11, nu = 0.8, the
    manufactured displacement (9e-4 cos(2 pi x) sin(2 pi y), 7e-4 sin(2 pi x)
    cos(2 pi y)) with its body force, prescribed on full-way bounce-back walls or
    as tractions on anti-bounce-back walls; ``steps_factor`` n^2 (+1) steps (the
    drivers: 3); L2 errors of the displacement and of the stress (scaled by
    1 / epsilon, epsilon = 1 / nx) relative to the exact fields, interior cells."""
    lam = E_EL * NU_EL / (1 - NU_EL ** 2)
    mu = 0.5 * E_EL / (1 + NU_EL)
    make = P.dirichlet_elasticity if walls == "dirichlet" else P.neumann_elasticity
    el = make(kind, n=n, lam=lam, mu=mu, precision=precision)
    f = el.run(steps_factor * n * n)
    eta, sigma = (_host(x) for x in el.update_macroscopic(f))
    nx, ny = el.nx, el.ny
    eta_ex = np.concatenate((P.sum_function(nx, ny, [[], [2, 2], [], []], lambda k, l: 9e-4, True),
                             P.sum_function(nx, ny, [[], [], [2, 2], []], lambda k, l: 7e-4, True)), -1)
    sig_ex = P._sigma_wall(nx, ny, lam, mu)
    sigma = sigma * nx  # / epsilon
    inner = (slice(1, -1), slice(1, -1))
    return {"eta": float(np.linalg.norm((eta - eta_ex)[inner]) / np.linalg.norm(eta_ex[inner])),
            "sigma": float(np.linalg.norm((sigma - sig_ex)[inner]) / np.linalg.norm(sig_ex[inner]))}


# ---------------------------------------------------------------------------
# periodic Biot, manufactured solution

def biot_periodic(kind, n, precision="f64/f64", coupling="centered", ne=None):
    """examples/poroelasticity/periodic.py:48-149, 174-259: E = 0.11, nu = 0.8,
    K = 0.1, alpha = 1, single grid, ``ne`` = n elasticity steps per physical
    step, n^2 physical steps (t_max + 1 steps are run), output every step; errors
    of the displacement, the stress (/ epsilon) and the pressure, accumulated over
    all steps and relative to the exact fields."""
    lam = E_EL * NU_EL / (1 - NU_EL ** 2)
    mu = 0.5 * E_EL / (1 + NU_EL)
    K = 0.1
    b = P.biot_periodic(kind, nx=n, levels=1, steps=ne or n, coupling=coupling, lam=lam, mu=mu, K=K)
    acc = []
    S = lambda kls, scale: P.sum_function(n, n, kls, scale)

    def output_data(**kw):
        t = kw["timestep"]
        Kd = (1.0 / P.mean_rate(b.darcy) - 0.5) / 3
        tf = 1 - np.exp(-8 * np.pi ** 2 * Kd * t / n ** 2)
        eta = np.array(kw["eta"])[0]
        sigma = np.array(kw["sigma"])[0] / b.epsilon
        p = np.array(kw["pressure"])[0]
        eta_ex = np.concatenate((S([[], [2, 2], [], []], lambda k, l: 4.5 * tf),
                                 S([[], [], [2, 2], []], lambda k, l: 3.5 * tf)), -1)
        sig_ex = np.concatenate((S([[], [], [], [2, 2]], lambda k, l: -np.pi * tf * (16 * lam + 18 * mu)),
                                 S([[2, 2], [], [], []], lambda k, l: np.pi * tf * 16 * mu),
                                 S([[], [], [], [2, 2]], lambda k, l: -np.pi * tf * (16 * lam + 14 * mu))), -1)
        p_ex = S([[], [], [], [2, 2]], lambda k, l: -np.pi / b.alpha * tf * (16 * lam + 30 * mu))
        acc.append([np.sum((eta - eta_ex) ** 2), np.sum(eta_ex ** 2), np.sum((sigma - sig_ex) ** 2),
                    np.sum(sig_ex ** 2), np.sum((p - p_ex) ** 2), np.sum(p_ex ** 2)])

    b.output_data = output_data
    b.ioRate = 1
    b.run(n * n)
    v = np.array(acc)
    return {"eta": _rel(v[:, 0], v[:, 1]), "sigma": _rel(v[:, 2], v[:, 3]), "p": _rel(v[:, 4], v[:, 5])}


# ---------------------------------------------------------------------------
# Terzaghi's consolidation

def _xcorner_fix(Base, remove):
    """``get_missing_indices`` of the Terzaghi driver's BC subclasses
    (Terzaghi.py:22-112): every direction listed, paired with its opposite; the
    directions of ``remove[(right_wall, top)]`` cleared from the boundary mask in
    place at that corner cell of a left / right wall; the directions outside the
    mask then zeroed (ADR_BB and ADR_AntiBB share the ADR table)."""

    class Fixed(Base):
        def get_missing_indices(self, boundaryMask):
            nbd = len(self.indices[0])
            q = self.lattice.q
            imissing = np.vstack([np.arange(q, dtype="uint8")] * nbd)
            iknown = np.vstack([np.asarray(self.lattice.opp_indices)] * nbd)
            for i in range(nbd):
                x, y = int(self.indices[0][i]), int(self.indices[1][i])
                for (right, top), dirs in remove.items():
                    if x == (self.nx - 2 if right else 1) and y == (self.ny - 1 if top else 0):
                        for d in dirs:
                            boundaryMask[i, d] = False
            imissing[~boundaryMask] = 0
            iknown[~boundaryMask] = 0
            return imissing, iknown

    Fixed.__name__ = "Fixed" + Base.__name__
    return Fixed


# Terzaghi.py:23-50 (ADR_BB), 53-80 (ADR_AntiBB) and 84-111 (ElasticityAntiBB):
# directions removed at the lower/upper corners of the left (False) / right (True) wall
TERZAGHI_ADR = {(False, False): (1, 4), (False, True): (2, 8), (True, False): (1, 7), (True, True): (2, 5)}
TERZAGHI_EL = {(False, False): (1, 5), (False, True): (3, 6), (True, False): (1, 4), (True, True): (3, 7)}


def terzaghi_setup(n):
    """The Terzaghi driver's non-dimensional parameters (Terzaghi.py:247-300): E = 1,
    nu = 0.8, c0 = 1, alpha = 1, L = 1, in units of L / n, eta0, T_tot / (4 n^2), p0."""
    E, nu, c0, alpha, L = 1.0, 0.8, 1.0, 1.0, 1.0
    lam = E * nu / (1 - nu ** 2)
    mu = 0.5 * E / (1 + nu)
    force = alpha / c0 + (lam + 2 * mu) / alpha
    Kpm = lam + 2 * mu + alpha ** 2 / c0
    p0 = alpha * force / c0 / Kpm
    eta0 = force * L / Kpm
    eta_inf = force * L / (lam + 2 * mu)
    kappa = Kpm / (lam + 2 * mu)
    c_f = kappa * (lam + 2 * mu) / c0 / Kpm
    T_tot = L ** 2 / c_f
    beta = 0.25
    uL, uH, uT, uP = L / n, eta0, T_tot * beta / n ** 2, p0
    return dict(lam=lam / uP, mu=mu / uP, alpha=alpha, epsilon=uL / uH, kappa=kappa * uT * uP / uL ** 2,
                force=force * uL / (uH * uP), eta_inf=eta_inf / uH, nt=int(T_tot / uT), dx=uL, dt=uT)


def terzaghi(kind, n, precision="f64/f64", n_exact=100, ne=None, multigrid=False):
    """examples/poroelasticity/Terzaghi.py: a column of n cells (+ two wall layers),
    fixed and impermeable on the left, loaded and drained on the right; 4 n^2
    physical steps, output every step; errors of the subsidence (mean displacement
    of the last fluid column) and of the pressure against the series solutions
    (``n_exact`` + 1 terms), accumulated over all steps and relative.  The driver's
    default is a single grid 4 cells high with ``ne`` = n elasticity steps per
    physical step; with ``multigrid`` (its ``-l -1``, the paper's sweep
    runTerzaghi.sh) the grid is n cells high, the elasticity sub-step ``ne`` V-cycles
    (default 1) over log2(n) levels with 4 smoothing steps, the walls kept on every
    level (Terzaghi.py:290-303, 316-323)."""
    m = P.modules(kind)
    s = terzaghi_setup(n)
    levels, ny = 1, 4
    if multigrid:
        levels = int(np.log2(n))
        while n % 2 ** (levels - 1) != 0:
            levels -= 1
        ny = n
        ne = ne or 1
    ADR_BB = _xcorner_fix(m["ADR_BB"], TERZAGHI_ADR)
    ADR_AntiBB = _xcorner_fix(m["ADR_ABB"], TERZAGHI_ADR)
    ElAntiBB = _xcorner_fix(m["EABB"], TERZAGHI_EL)
    EBBF = m["EBBF"]
    load = -s["force"]

    class Darcy(m["Darcy"]):
        def initialize_macroscopic_fields(self):
            return np.zeros((self.nx, self.ny, 1)), np.zeros((self.nx, self.ny, 2))

        def set_boundary_conditions(self):
            self.BCs.append(ADR_BB(tuple(self.boundingBoxIndices["left"].T), self.gridInfo, self.precisionPolicy))
            self.BCs.append(ADR_AntiBB(tuple(self.boundingBoxIndices["right"].T), self.gridInfo, self.precisionPolicy))

    class El(m["Elasticity"]):
        def initialize_macroscopic_fields(self):
            return np.zeros((self.nx, self.ny, 2)), np.zeros((self.nx, self.ny, 3))

        def set_boundary_conditions(self):
            self.BCs.append(EBBF(tuple(self.boundingBoxIndices["left"].T), self.gridInfo, self.precisionPolicy))
            force = np.stack((np.full((self.ny,), load), np.zeros((self.ny,))), axis=-1)
            self.BCs.append(ElAntiBB(tuple(self.boundingBoxIndices["right"].T), self.gridInfo, self.precisionPolicy,
                                     force))
            for level in range(1, self.levels):
                gi = self.grid_info(level)
                bbl = self.bounding_box_indices(level)
                self.BCs_MG[level] = [EBBF(tuple(bbl["left"].T), gi, self.precisionPolicy),
                                      ElAntiBB(tuple(bbl["right"].T), gi, self.precisionPolicy)]

    acc = []
    j = np.arange(n_exact + 1)

    class Terzaghi(m["Poro"]):
        def __init__(self, **kw):
            super().__init__(El, Darcy, **kw)

        def output_data(self, **kw):
            t = kw["timestep"]
            eta = np.array(kw["eta"])[0]
            p = np.array(kw["pressure"])[0]
            nx, ny = p.shape[0], p.shape[1]
            if t > 0:
                tc = -0.25 * np.pi ** 2 * t * s["dt"]
                x = np.arange(nx)[:, None]
                series = ((-1.0) ** j / (2 * j + 1))[None, :] * np.cos((2 * j + 1)[None, :] * 0.5 * np.pi * s["dx"]
                                                                       * (x - 0.5)) * np.exp((2 * j + 1)[None, :] ** 2 * tc)
                p_ex = np.broadcast_to((4.0 / np.pi * series.sum(1))[:, None, None], (nx, ny, 1))
                sub = s["eta_inf"] + (1.0 - s["eta_inf"]) * 8.0 / np.pi ** 2 * np.sum(
                    np.exp((2 * j + 1) ** 2 * tc) / (2 * j + 1) ** 2)
            else:
                p_ex = np.ones((nx, ny, 1))
                sub = 1.0
            acc.append([(np.mean(eta[-2, :, 0]) + sub) ** 2, sub ** 2,
                        np.sum((p - p_ex)[1:-1] ** 2), np.sum(p_ex[1:-1] ** 2)])

    sim = Terzaghi(**{"lattice_elasticity": m["D2Q8"](precision), "lattice_darcy": m["D2Q9"](precision),
                      "elasticity_steps": ne or n, "elasticity_MG_levels": levels,
                      "MG_BCs": {"left": True, "right": True}, "MG_n_smoothing": 4, "coupling_method": "centered",
                      "alpha": s["alpha"], "lambda": s["lam"], "mu": s["mu"], "epsilon": s["epsilon"],
                      "omega": 1.0 / (3 * s["kappa"] + 0.5), "nx": n + 2, "ny": ny, "nz": 0, "precision": precision,
                      "io_rate": 1, "print_info_rate": 0})
    sim.run(s["nt"])
    v = np.array(acc)
    return {"subsidence": _rel(v[:, 0], v[:, 1]), "p": _rel(v[:, 2], v[:, 3])}
