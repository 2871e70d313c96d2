"""Problem definitions shared by the GPU solvers and the CPU oracle.

Each factory builds the same configuration on either implementation: ``kind="gpu"``
uses the drop-in classes of ``xlb-biot_amd/src``, ``kind="oracle"`` the CPU
restatement in ``oracle/``.  The set-ups follow the reference drivers
(``examples/elasticity/*_MG.py``, ``examples/darcy/*.py``,
``examples/poroelasticity/{periodic,2D}.py``).
"""
import numpy as np


def modules(kind):
    if kind == "gpu":
        from src import darcy as D
        from src import elasticity as E
        from src import lattice as L
        from src import poroelasticity as P
        return dict(Elasticity=E.LBMElasticity, Darcy=D.AdvectionDiffusionReactionBGK, Poro=P.LBMPoroelasticity,
                    D2Q8=L.LatticeD2Q8, D2Q9=L.LatticeD2Q9, EBBF=E.ElasticityBounceBackFullway,
                    EBB=E.ElasticityBounceBack, EABB=E.ElasticityAntiBounceBack, ADR_ABB=D.ADR_AntiBounceBack,
                    ADR_BB=D.ADR_Neumann_BounceBack, ADR_EQ=D.ADR_EquilibriumBC)
    from oracle import bc as B
    from oracle import darcy as OD
    from oracle import elasticity as OE
    from oracle import lattice as OL
    from oracle import poroelasticity as OP
    return dict(Elasticity=OE.OracleElasticity, Darcy=OD.OracleDarcy, Poro=OP.OraclePoroelasticity,
                D2Q8=lambda precision="f64/f64": OL.D2Q8(), D2Q9=lambda precision="f64/f64": OL.D2Q9(),
                EBBF=B.ElasticityBounceBackFullway, EBB=B.ElasticityBounceBack, EABB=B.ElasticityAntiBounceBack,
                ADR_ABB=B.ADR_AntiBounceBack, ADR_BB=B.ADR_Neumann_BounceBack, ADR_EQ=B.ADR_EquilibriumBC)


def sum_function(nx, ny, kls, scale=lambda k, l: 1.0, walls=False):
    """``sumFunction`` of the reference drivers: periodic cell centres ``(x+1/2)/nx``
    or, with ``walls``, ``(x-1/2)/(nx-2)`` (one BC layer per side)."""
    x = np.arange(nx)[:, None, None].astype(np.float64)
    y = np.arange(ny)[None, :, None].astype(np.float64)
    if walls:
        X, Y = (x - 0.5) / (nx - 2), (y - 0.5) / (ny - 2)
    else:
        X, Y = (x + 0.5) / nx, (y + 0.5) / ny
    val = np.zeros((nx, ny, 1))
    for j, (f, g) in enumerate([(np.cos, np.cos), (np.cos, np.sin), (np.sin, np.cos), (np.sin, np.sin)]):
        for i in range(len(kls[j]) // 2):
            k, l = kls[j][2 * i], kls[j][2 * i + 1]
            val = val + scale(k, l) * f(k * np.pi * X) * g(l * np.pi * Y)
    return val


def manufactured_c1(nx):
    """``(sin 2pi x sin 2pi y, cos 2pi x cos 2pi y)`` at cell centres."""
    X = (np.arange(nx) + 0.5) / nx
    a, b = 2 * np.pi * X[:, None], 2 * np.pi * X[None, :]
    return np.stack([np.sin(a) * np.sin(b), np.cos(a) * np.cos(b)], -1)


def lognormal_field(nx, ny, variance, corr_len, seed, mean=0.0):
    """Log-normal field ``exp(g)``, g a Gaussian random field with exponential
    covariance ``variance * exp(-r / corr_len)``, by circulant embedding on a
    doubled torus (NumPy PCG64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Nx, Ny = 2 * nx, 2 * ny
    dx = np.minimum(np.arange(Nx), Nx - np.arange(Nx))[:, None]
    dy = np.minimum(np.arange(Ny), Ny - np.arange(Ny))[None, :]
    cov = variance * np.exp(-np.sqrt(dx ** 2 + dy ** 2) / corr_len)
    lam = np.real(np.fft.fft2(cov))
    lam[lam < 0] = 0.0
    xi = rng.standard_normal((Nx, Ny)) + 1j * rng.standard_normal((Nx, Ny))
    g = np.real(np.fft.ifft2(np.sqrt(lam) * xi)) * np.sqrt(Nx * Ny)
    return np.exp(mean + g[:nx, :ny])


# ---------------------------------------------------------------------------
# elasticity

def periodic_elasticity(kind, nx=64, lam=1.0, mu=1.0, precision="f64/f64", levels=1, k_smooth=2, gamma=0.75):
    """C1: periodic D2Q8, manufactured body force, zero initial populations."""
    m = modules(kind)

    class C1(m["Elasticity"]):
        def get_force(self):
            eps = 1.0 / self.nx
            return eps ** 2 * self.mu * manufactured_c1(self.nx)

        def initialize_macroscopic_fields(self):
            return None, None

    return C1(**{"lambda": lam, "mu": mu, "nx": nx, "ny": nx, "nz": 0, "precision": precision,
                 "lattice": m["D2Q8"](precision), "MG_levels": levels, "MG_n_smoothing": k_smooth,
                 "MG_relaxation": gamma, "print_info_rate": 0})


def _sigma_wall(nx, ny, lam, mu):
    return np.concatenate((sum_function(nx, ny, [[], [], [], [2, 2]], lambda k, l: np.pi * (-3.2e-3 * lam - 3.6e-3 * mu), True),
                           sum_function(nx, ny, [[2, 2], [], [], []], lambda k, l: np.pi * 3.2e-3 * mu, True),
                           sum_function(nx, ny, [[], [], [], [2, 2]], lambda k, l: np.pi * (-3.2e-3 * lam - 2.8e-3 * mu), True)),
                          axis=-1)


def _wall_force(self):
    eps = 1.0 / (self.nx - 2)
    f = np.concatenate((sum_function(self.nx, self.ny, [[], [2, 2], [], []],
                                     lambda k, l: eps ** 2 * np.pi ** 2 * (1.36e-2 * self.mu + 6.4e-3 * self.lamda), True),
                        sum_function(self.nx, self.ny, [[], [], [2, 2], []],
                                     lambda k, l: eps ** 2 * np.pi ** 2 * (1.20e-2 * self.mu + 6.4e-3 * self.lamda), True)),
                       axis=-1)
    f[0, ...] = 0.0
    f[1, ...] = 0.0
    f[:, 0, ...] = 0.0
    f[:, 1, ...] = 0.0
    return f


def neumann_elasticity(kind, n=16, levels=1, lam=0.5, mu=0.4, precision="f64/f64", k_smooth=2):
    """examples/elasticity/neumann_MG.py: four traction (anti-bounce-back) walls."""
    m = modules(kind)
    EABB = m["EABB"]

    class Neumann(m["Elasticity"]):
        def get_force(self):
            return _wall_force(self)

        def set_boundary_conditions(self):
            sw = _sigma_wall(self.nx, self.ny, self.lamda, self.mu)

            def T_wall(nrm):
                eps = 1.0 / (self.nx - 2)
                return eps * np.stack((sw[..., 0] * nrm[0] + sw[..., 1] * nrm[1],
                                       sw[..., 1] * nrm[0] + sw[..., 2] * nrm[1]), axis=-1)
            bb = self.boundingBoxIndices
            self.BCs.append(EABB(tuple(bb["left"].T), self.gridInfo, self.precisionPolicy, T_wall([-1, 0])[-1, 1:-1]))
            self.BCs.append(EABB(tuple(bb["right"].T), self.gridInfo, self.precisionPolicy, T_wall([1, 0])[0, 1:-1]))
            self.BCs.append(EABB(tuple(bb["bottom"].T), self.gridInfo, self.precisionPolicy, T_wall([0, -1])[1:-1, -1]))
            self.BCs.append(EABB(tuple(bb["top"].T), self.gridInfo, self.precisionPolicy, T_wall([0, 1])[1:-1, 0]))
            for level in range(1, self.levels):
                gi = self.grid_info(level)
                bbl = self.bounding_box_indices(level)
                self.BCs_MG[level] = [EABB(tuple(bbl[s].T), gi, self.precisionPolicy)
                                      for s in ("left", "right", "bottom", "top")]

    return Neumann(**{"lambda": lam, "mu": mu, "nx": n + 2, "ny": n + 2, "nz": 0, "precision": precision,
                      "lattice": m["D2Q8"](precision), "MG_levels": levels, "MG_n_smoothing": k_smooth,
                      "MG_BCs": {"top": True, "bottom": True, "left": True, "right": True}, "print_info_rate": 0})


def dirichlet_elasticity(kind, n=16, levels=1, lam=0.5, mu=0.4, precision="f64/f64", k_smooth=2):
    """examples/elasticity/dirichlet_MG.py: four full-way bounce-back walls with a
    prescribed displacement."""
    m = modules(kind)
    EBBF = m["EBBF"]

    class Dirichlet(m["Elasticity"]):
        def get_force(self):
            return _wall_force(self)

        def set_boundary_conditions(self):
            ew = np.concatenate((sum_function(self.nx, self.ny, [[], [2, 2], [], []], lambda k, l: 9e-4, True),
                                 sum_function(self.nx, self.ny, [[], [], [2, 2], []], lambda k, l: 7e-4, True)), -1)
            bb = self.boundingBoxIndices
            self.BCs.append(EBBF(tuple(bb["left"].T), self.gridInfo, self.precisionPolicy, ew[-1, ...]))
            self.BCs.append(EBBF(tuple(bb["right"].T), self.gridInfo, self.precisionPolicy, ew[0, ...]))
            self.BCs.append(EBBF(tuple(bb["bottom"].T), self.gridInfo, self.precisionPolicy, ew[:, -1, ...]))
            self.BCs.append(EBBF(tuple(bb["top"].T), self.gridInfo, self.precisionPolicy, ew[:, 0, ...]))
            for level in range(1, self.levels):
                gi = self.grid_info(level)
                bbl = self.bounding_box_indices(level)
                self.BCs_MG[level] = [EBBF(tuple(bbl[s].T), gi, self.precisionPolicy)
                                      for s in ("left", "right", "bottom", "top")]

    return Dirichlet(**{"lambda": lam, "mu": mu, "nx": n + 2, "ny": n + 2, "nz": 0, "precision": precision,
                        "lattice": m["D2Q8"](precision), "MG_levels": levels, "MG_n_smoothing": k_smooth,
                        "MG_BCs": {"top": True, "bottom": True, "left": True, "right": True},
                        "print_info_rate": 0})


def obstacle_elasticity(kind, n=16, lam=0.5, mu=0.4, precision="f64/f64", levels=1, k_smooth=2):
    """The four full-way Dirichlet walls of ``dirichlet_elasticity`` around an interior
    L-shaped solid obstacle held at a prescribed displacement (halfway bounce-back on
    solid cells: the fluid cells around it take their missing directions from the grid
    mask, diagonals included).  With ``levels`` > 1 every coarse level carries its own
    walls and the obstacle's coarse voxels (interior cell i -> i // 2^level), both
    homogeneous, as ``BCs_MG`` (elasticity.py:396-436)."""
    m = modules(kind)
    EBBF, EBB = m["EBBF"], m["EBB"]
    block = [(x, y) for x in range(6, 9) for y in range(6, 12)] + [(x, 6) for x in range(9, 12)]

    class Obstacle(m["Elasticity"]):
        def get_force(self):
            return _wall_force(self)

        def set_boundary_conditions(self):
            ew = np.concatenate((sum_function(self.nx, self.ny, [[], [2, 2], [], []], lambda k, l: 9e-4, True),
                                 sum_function(self.nx, self.ny, [[], [], [2, 2], []], lambda k, l: 7e-4, True)), -1)
            bb = self.boundingBoxIndices
            self.BCs.append(EBBF(tuple(bb["left"].T), self.gridInfo, self.precisionPolicy, ew[-1, ...]))
            self.BCs.append(EBBF(tuple(bb["right"].T), self.gridInfo, self.precisionPolicy, ew[0, ...]))
            self.BCs.append(EBBF(tuple(bb["bottom"].T), self.gridInfo, self.precisionPolicy, ew[:, -1, ...]))
            self.BCs.append(EBBF(tuple(bb["top"].T), self.gridInfo, self.precisionPolicy, ew[:, 0, ...]))
            # one displacement vector for the whole obstacle (the per-cell form indexes the
            # given rows with the shifted cells' positions, elasticity.py:1604-1608)
            self.BCs.append(EBB(tuple(np.array(block).T), self.gridInfo, self.precisionPolicy,
                                np.array([2e-4, -1e-4])))
            for level in range(1, self.levels):
                gi = self.grid_info(level)
                bbl = self.bounding_box_indices(level)
                coarse = sorted({((x - 1) // 2 ** level + 1, (y - 1) // 2 ** level + 1) for x, y in block})
                self.BCs_MG[level] = [EBBF(tuple(bbl[s].T), gi, self.precisionPolicy)
                                      for s in ("left", "right", "bottom", "top")]
                self.BCs_MG[level].append(EBB(tuple(np.array(coarse).T), gi, self.precisionPolicy))

    kw = {"lambda": lam, "mu": mu, "nx": n + 2, "ny": n + 2, "nz": 0, "precision": precision,
          "lattice": m["D2Q8"](precision), "print_info_rate": 0}
    if levels > 1:
        kw.update({"MG_levels": levels, "MG_n_smoothing": k_smooth,
                   "MG_BCs": {"top": True, "bottom": True, "left": True, "right": True}})
    return Obstacle(**kw)


def mixed_halfway_elasticity(kind, nx=18, ny=12, precision="f64/f64", lam=1.0, mu=0.7):
    """Halfway Dirichlet (``ElasticityBounceBack``) on the left, traction on the right,
    periodic in y: exercises the duplicate-index scatters of the halfway bounce-back
    and the y-halo marked solid next to a periodic face."""
    m = modules(kind)
    EBB, EABB = m["EBB"], m["EABB"]

    class Mixed(m["Elasticity"]):
        def get_force(self):
            x = (np.arange(self.nx)[:, None] + 0.5) / self.nx
            y = (np.arange(self.ny)[None, :] + 0.5) / self.ny
            return 1e-3 * np.stack([np.sin(2 * np.pi * y) + 0 * x, np.cos(np.pi * x) + 0 * y], -1)

        def set_boundary_conditions(self):
            bb = self.boundingBoxIndices
            self.BCs.append(EBB(tuple(bb["left"].T), self.gridInfo, self.precisionPolicy, np.array([1e-3, -2e-3])))
            T = np.stack([np.linspace(1e-3, 2e-3, self.ny), np.linspace(-5e-4, 5e-4, self.ny)], -1)
            self.BCs.append(EABB(tuple(bb["right"].T), self.gridInfo, self.precisionPolicy, T))

    return Mixed(**{"lambda": lam, "mu": mu, "nx": nx, "ny": ny, "nz": 0, "precision": precision,
                    "lattice": m["D2Q8"](precision), "print_info_rate": 0})


# ---------------------------------------------------------------------------
# Darcy

def darcy_c2(kind, n=32, precision="f64/f64", heterogeneous=True, seed=0, variance=1.0, corr=16.0, source=False):
    """C2: log-normal permeability, p=1 on the left, p=0 on the right (anti-bounce-back),
    no flux at the bottom and top (bounce-back); one ghost layer per wall."""
    m = modules(kind)
    ADR_ABB, ADR_BB = m["ADR_ABB"], m["ADR_BB"]
    nx = ny = n + 2

    class C2(m["Darcy"]):
        def set_boundary_conditions(self):
            bb = self.boundingBoxIndices
            self.BCs.append(ADR_ABB(tuple(bb["left"].T), self.gridInfo, self.precisionPolicy, 1.0))
            self.BCs.append(ADR_ABB(tuple(bb["right"].T), self.gridInfo, self.precisionPolicy, 0.0))
            self.BCs.append(ADR_BB(tuple(bb["bottom"].T), self.gridInfo, self.precisionPolicy))
            self.BCs.append(ADR_BB(tuple(bb["top"].T), self.gridInfo, self.precisionPolicy))

        def get_source(self):
            if not source:
                return None
            return 1e-3 * sum_function(self.nx, self.ny, [[], [], [], [2, 2]], walls=True)

        def initialize_macroscopic_fields(self):
            x = (np.arange(self.nx)[:, None, None] - 0.5) / (self.nx - 2)
            return np.clip(1.0 - x, 0.0, 1.0) * np.ones((1, self.ny, 1)), np.zeros((self.nx, self.ny, 2))

    if heterogeneous:
        K = lognormal_field(nx, ny, variance, corr, seed)
        omega = 1.0 / (3.0 * K + 0.5)
        omega = omega.reshape(nx, ny, 1)
    else:
        omega = 1.0 / (3.0 * 1.0 + 0.5)
    return C2(**{"omega": omega, "nx": nx, "ny": ny, "nz": 0, "precision": precision,
                 "lattice": m["D2Q9"](precision), "print_info_rate": 0})


def darcy_obstacle(kind, n=14, precision="f64/f64"):
    """The C2 walls around an interior solid obstacle: an L-shaped voxel set whose
    surface holds p = 0.3 (anti-bounce-back on solid cells, so the fluid cells around
    it take their missing directions from the grid mask, diagonals included)."""
    m = modules(kind)
    ADR_ABB, ADR_BB = m["ADR_ABB"], m["ADR_BB"]
    nx = ny = n + 2

    class Obstacle(m["Darcy"]):
        def set_boundary_conditions(self):
            bb = self.boundingBoxIndices
            self.BCs.append(ADR_ABB(tuple(bb["left"].T), self.gridInfo, self.precisionPolicy, 1.0))
            self.BCs.append(ADR_ABB(tuple(bb["right"].T), self.gridInfo, self.precisionPolicy, 0.0))
            self.BCs.append(ADR_BB(tuple(bb["bottom"].T), self.gridInfo, self.precisionPolicy))
            self.BCs.append(ADR_BB(tuple(bb["top"].T), self.gridInfo, self.precisionPolicy))
            block = [(x, y) for x in range(6, 9) for y in range(5, 11)] + [(x, 5) for x in range(9, 11)]
            self.BCs.append(ADR_ABB(tuple(np.array(block).T), self.gridInfo, self.precisionPolicy, 0.3))

        def initialize_macroscopic_fields(self):
            x = (np.arange(self.nx)[:, None, None] - 0.5) / (self.nx - 2)
            return np.clip(1.0 - x, 0.0, 1.0) * np.ones((1, self.ny, 1)), np.zeros((self.nx, self.ny, 2))

    K = lognormal_field(nx, ny, 1.0, 4.0, 2)
    return Obstacle(**{"omega": (1.0 / (3.0 * K + 0.5)).reshape(nx, ny, 1), "nx": nx, "ny": ny, "nz": 0,
                       "precision": precision, "lattice": m["D2Q9"](precision), "print_info_rate": 0})


def darcy_equilibrium(kind, n=16, precision="f64/f64"):
    """examples/darcy/dirichlet_eq.py-like: equilibrium walls with a manufactured source."""
    m = modules(kind)
    ADR_EQ = m["ADR_EQ"]
    nx = ny = n + 2

    class Eq(m["Darcy"]):
        def set_boundary_conditions(self):
            walls = np.concatenate((self.boundingBoxIndices["left"], self.boundingBoxIndices["right"],
                                    self.boundingBoxIndices["top"], self.boundingBoxIndices["bottom"]))
            self.BCs.append(ADR_EQ(tuple(walls.T), self.gridInfo, self.precisionPolicy, 0.5))

        def get_source(self):
            K = (1.0 / self.omega - 0.5) / 3
            s = sum_function(self.nx, self.ny, [[], [], [], [2, 4, 4, 2]],
                             lambda k, l: 0.5 * K * (k ** 2 + l ** 2) * np.pi ** 2 / (self.nx - 2) ** 2, True)
            s[(0, -1), :] = 0
            s[:, (0, -1)] = 0
            return s

        def initialize_macroscopic_fields(self):
            p = sum_function(self.nx, self.ny, [[], [], [], [2, 4, 4, 2]], walls=True)
            return p, np.zeros((self.nx, self.ny, 2))

    return Eq(**{"omega": 1.0 / (3 * 0.1 + 0.5), "nx": nx, "ny": ny, "nz": 0, "precision": precision,
                 "lattice": m["D2Q9"](precision), "print_info_rate": 0})


# ---------------------------------------------------------------------------
# Biot

def mean_rate(darcy):
    """The Darcy solver's mean relaxation rate: the scalar ``omega``, or the mean of a
    per-cell ``omega`` field (the solver keeps the mean in ``omega``, the oracle keeps
    the field itself)."""
    om = darcy.omega
    if isinstance(om, float):
        return om
    return float(np.asarray(om.detach().cpu() if hasattr(om, "detach") else om, dtype=np.float64).mean())


def step_rel(out, ref, against_state=()):
    """Relative error of each output of a coupled Biot step against the oracle's:
    ``max|a - b| / max|b|``.  An output that vanishes in exact arithmetic (``max|b|``
    below 1e-8 of the populations' scale ``S = max(max|f_el|, max|f_darcy|)``) is
    measured against ``S`` instead: its entries are round-off of sums of
    populations.  So are the outputs listed in ``against_state`` (indices into the
    step's 7-tuple), such as the divergence rate (output 4), which is formed from
    the difference of two consecutive states and so carries their rounding."""
    import torch
    cpu = [torch.as_tensor(x).detach().to("cpu", torch.float64) for x in out]
    ref = [torch.as_tensor(x).detach().to("cpu", torch.float64) for x in ref]
    S = max(ref[0].abs().max().item(), ref[1].abs().max().item())
    errs = []
    for k, (a, b) in enumerate(zip(cpu, ref)):
        assert a.shape == b.shape, (a.shape, b.shape)
        m = b.abs().max().item()
        den = m if m > 1e-8 * S and k not in against_state else S
        errs.append((a - b).abs().max().item() / (den if den > 0 else 1.0))
    return errs


def biot_periodic(kind, nx=32, levels=1, steps=1, precision="f64/f64", coupling="centered", alt=False,
                  lam=1.0, mu=1.0, alpha=1.0, K=1.0, K_field=None, k_smooth=None, gamma=None, device_hooks=False):
    """examples/poroelasticity/periodic.py: manufactured time-dependent solution.
    ``K_field`` (nx, nx) gives a heterogeneous permeability (per-cell Darcy rate
    1/(3K+1/2)); the manufactured forcing then uses the mean rate.  With
    ``device_hooks`` the hooks keep the (time-independent) spatial patterns on the
    GPU and return SoA-backed views scaled by the time factor."""
    m = modules(kind)

    class PDarcy(m["Darcy"]):
        def initialize_macroscopic_fields(self):
            return np.zeros((self.nx, self.ny, 1)), np.zeros((self.nx, self.ny, 2))

    class PEl(m["Elasticity"]):
        def initialize_macroscopic_fields(self):
            return np.zeros((self.nx, self.ny, 2)), np.zeros((self.nx, self.ny, 3))

    base = m["Poro"]

    class Periodic(base):
        def __init__(self, **kw):
            super().__init__(PEl, PDarcy, **kw)

        def _pattern(self, name, make):
            import torch
            pats = self.__dict__.setdefault("_patterns", {})
            if name not in pats:
                pats[name] = torch.as_tensor(make()[..., 0], dtype=torch.float64, device="cuda")
            return pats[name]

        def get_elastic_force(self, timestep):
            Kd = (1.0 / mean_rate(self.darcy) - 0.5) / 3
            tf = 1 - np.exp(-8 * np.pi ** 2 * Kd * timestep / self.elasticity.nx ** 2)
            if device_hooks:
                import torch
                pat = self._pattern("force", lambda: sum_function(
                    self.elasticity.nx, self.elasticity.ny, [[], [2, 2], [], []],
                    lambda k, l: self.epsilon ** 2 * np.pi ** 2 * 8 * self.elasticity.mu))
                out = torch.zeros((2,) + tuple(pat.shape), dtype=torch.float64, device="cuda")
                out[0] = pat * tf
                return out.permute(1, 2, 0)
            fx = sum_function(self.elasticity.nx, self.elasticity.ny, [[], [2, 2], [], []],
                              lambda k, l: self.epsilon ** 2 * tf * np.pi ** 2 * 8 * self.elasticity.mu)
            return np.concatenate((fx, np.zeros_like(fx)), -1)

        def get_darcy_source(self, timestep):
            Kd = (1.0 / mean_rate(self.darcy) - 0.5) / 3
            tf = np.exp(-8 * np.pi ** 2 * Kd * timestep / self.elasticity.nx ** 2)
            factor = -np.pi ** 3 * Kd / self.darcy.nx ** 2 * (
                128 * self.alpha * tf + (128 * self.elasticity.lamda + 240 * self.elasticity.mu) / self.alpha)
            if device_hooks:
                pat = self._pattern("source", lambda: sum_function(self.darcy.nx, self.darcy.ny,
                                                                   [[], [], [], [2, 2]], lambda k, l: 1.0))
                return (pat * factor).unsqueeze(-1)
            return sum_function(self.darcy.nx, self.darcy.ny, [[], [], [], [2, 2]], lambda k, l: factor)

    mg = {}
    if k_smooth is not None:
        mg["MG_n_smoothing"] = k_smooth
    if gamma is not None:
        mg["MG_relaxation"] = gamma
    return Periodic(**mg, **{"lattice_elasticity": m["D2Q8"](precision), "lattice_darcy": m["D2Q9"](precision),
                       "elasticity_steps": steps, "elasticity_MG_levels": levels, "coupling_method": coupling,
                       "alternative_gradient": alt, "alpha": alpha, "lambda": lam, "mu": mu,
                       "omega": (1.0 / (3 * K + 0.5) if K_field is None
                                 else (1.0 / (3.0 * K_field + 0.5)).reshape(nx, nx, 1)),
                       "nx": nx, "ny": nx, "nz": 0, "precision": precision, "print_info_rate": 0})


def c5_specs(n_inst=64, seed=3):
    """Batched Biot instances: square sizes drawn from {64, 128, 256, 512}, lambda and
    mu ~ U[0.5, 2], log-permeability mean ~ U[-1, 1] (NumPy PCG64, one seed); each
    instance's heterogeneity is a GRF of variance 1 and correlation length 16."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = rng.choice([64, 128, 256, 512], size=n_inst)
    lam = rng.uniform(0.5, 2.0, n_inst)
    mu = rng.uniform(0.5, 2.0, n_inst)
    logk = rng.uniform(-1.0, 1.0, n_inst)
    return [dict(n=int(sizes[i]), lam=float(lam[i]), mu=float(mu[i]), logk_mean=float(logk[i]),
                 seed=1000 * seed + i) for i in range(n_inst)]


def c5_permeability(spec):
    return lognormal_field(spec["n"], spec["n"], 1.0, 16.0, spec["seed"], mean=spec["logk_mean"])


def biot_c5(kind, spec, levels, tol=1e-8, max_cycles=20000, precision="f64/f64", K=None, k_smooth=None,
            gamma=None):
    """One batched-Biot instance: the periodic manufactured Biot problem with the
    spec's size, Lame parameters and log-normal permeability, multigrid elasticity
    iterated to ``tol`` (relative to the residual of the zero state) each physical
    step.  The pressure force is the lattice gradient of p (the reference's
    ``alternative_gradient``): with a heterogeneous permeability the default form
    -3 omega(x) sum_i c_i f_i has a nonzero mean, a net force on the periodic body,
    for which the elastic problem has no steady state."""
    n = spec["n"]
    K = c5_permeability(spec) if K is None else K
    p = biot_periodic(kind, nx=n, levels=levels, lam=spec["lam"], mu=spec["mu"], precision=precision, K_field=K,
                      k_smooth=k_smooth, gamma=gamma, device_hooks=kind == "gpu", alt=True)
    p.elasticityTol = tol
    p.elasticityMaxSteps = max_cycles
    p.elasticityTolScale = "rhs"
    return p


def _corner_fix(Base, remove):
    """Subclass overriding ``get_missing_indices`` to drop corner directions in place,
    like the drivers' ADR_BB / ADR_AntiBB / ElasticityAntiBB (2D.py:20-110)."""

    class Fixed(Base):
        def get_missing_indices(self, boundaryMask):
            nbd = len(self.indices[0])
            q = self.lattice.q
            imissing = np.vstack([np.arange(q, dtype="uint8")] * nbd)
            iknown = np.vstack([np.asarray(self.lattice.opp_indices)] * nbd)
            for i in range(nbd):
                x, y = self.indices[0][i], self.indices[1][i]
                for (cx, cy), dirs in remove.items():
                    X = {0: 0, 1: self.nx - 1}[cx]
                    Y = {0: 1, 1: self.ny - 2}[cy]
                    if x == X and y == Y:
                        for d in dirs:
                            boundaryMask[i, d] = False
            imissing[~boundaryMask] = 0
            iknown[~boundaryMask] = 0
            return imissing, iknown

    Fixed.__name__ = "Fixed" + Base.__name__
    return Fixed


ADR_CORNERS = {(0, 0): (3, 5), (0, 1): (3, 7), (1, 0): (6, 8), (1, 1): (4, 6)}
EL_CORNERS = {(0, 0): (0, 7), (0, 1): (0, 4), (1, 0): (2, 6), (1, 1): (2, 5)}


def biot_consolidation(kind, nx=32, levels=2, steps=2, precision="f64/f64", k_smooth=4, lam=2.0, mu=1.0,
                       alpha=1.0, K=0.5, load=1e-3, coupling="centered", lambda_field=None, mu_field=None,
                       omega_field=None):
    """C4-like 2D consolidation (examples/poroelasticity/2D.py): fixed bottom
    (full-way bounce-back), traction at the top (corner-fixed anti-bounce-back);
    undrained bottom, drained top; periodic in x; ``ny = nx + 2``.  Optional
    per-cell Lame fields and Darcy rate ``(nx, ny)`` (the oracle takes tensors,
    which may require grad)."""
    m = modules(kind)
    ADR_BB = _corner_fix(m["ADR_BB"], ADR_CORNERS)
    ADR_AntiBB = _corner_fix(m["ADR_ABB"], ADR_CORNERS)
    ElAntiBB = _corner_fix(m["EABB"], EL_CORNERS)
    EBBF = m["EBBF"]

    class Darcy(m["Darcy"]):
        def initialize_macroscopic_fields(self):
            return np.zeros((self.nx, self.ny, 1)), np.zeros((self.nx, self.ny, 2))

        def set_boundary_conditions(self):
            self.BCs.append(ADR_BB(tuple(self.boundingBoxIndices["bottom"].T), self.gridInfo, self.precisionPolicy))
            self.BCs.append(ADR_AntiBB(tuple(self.boundingBoxIndices["top"].T), self.gridInfo, self.precisionPolicy))

    class El(m["Elasticity"]):
        def initialize_macroscopic_fields(self):
            return np.zeros((self.nx, self.ny, 2)), np.zeros((self.nx, self.ny, 3))

        def set_boundary_conditions(self):
            self.BCs.append(EBBF(tuple(self.boundingBoxIndices["bottom"].T), self.gridInfo, self.precisionPolicy))
            force = np.fromfunction(lambda x, d: (d == 1) * (1 - np.cos(np.pi * ((2 * x + 1) / self.nx + 0.5))) * -load,
                                    (self.nx, 2))
            self.BCs.append(ElAntiBB(tuple(self.boundingBoxIndices["top"].T), self.gridInfo, self.precisionPolicy, force))
            for level in range(1, self.levels):
                gi = self.grid_info(level)
                bbl = self.bounding_box_indices(level)
                self.BCs_MG[level] = [EBBF(tuple(bbl["bottom"].T), gi, self.precisionPolicy),
                                      ElAntiBB(tuple(bbl["top"].T), gi, self.precisionPolicy)]

    class Consolidation(m["Poro"]):
        def __init__(self, **kw):
            super().__init__(El, Darcy, **kw)

    kw = {"lattice_elasticity": m["D2Q8"](precision), "lattice_darcy": m["D2Q9"](precision),
          "elasticity_steps": steps, "elasticity_MG_levels": levels,
          "MG_BCs": {"bottom": True, "top": True}, "MG_n_smoothing": k_smooth,
          "coupling_method": coupling, "alpha": alpha, "lambda": lam, "mu": mu,
          "epsilon": 1.0 / nx, "omega": 1.0 / (3 * K + 0.5), "nx": nx, "ny": nx + 2, "nz": 0,
          "precision": precision, "print_info_rate": 0}
    if lambda_field is not None:
        kw["lambda_field"], kw["mu_field"] = lambda_field, mu_field
    if omega_field is not None:
        kw["omega"] = omega_field
    return Consolidation(**kw)


def consolidation_elasticity(kind, nx=16, lam=2.0, mu=1.0, precision="f64/f64", load=1e-3, lambda_field=None,
                             mu_field=None):
    """The elastic half of the 2-D consolidation set-up (examples/poroelasticity/2D.py):
    fixed bottom (full-way bounce-back), traction at the top (corner-fixed
    anti-bounce-back), periodic in x, ``ny = nx + 2``; single grid.  Optional per-cell
    Lame fields (the oracle takes tensors, which may require grad)."""
    m = modules(kind)
    ElAntiBB = _corner_fix(m["EABB"], EL_CORNERS)
    EBBF = m["EBBF"]

    class El(m["Elasticity"]):
        def initialize_macroscopic_fields(self):
            return np.zeros((self.nx, self.ny, 2)), np.zeros((self.nx, self.ny, 3))

        def set_boundary_conditions(self):
            self.BCs.append(EBBF(tuple(self.boundingBoxIndices["bottom"].T), self.gridInfo, self.precisionPolicy))
            force = np.fromfunction(lambda x, d: (d == 1) * (1 - np.cos(np.pi * ((2 * x + 1) / self.nx + 0.5))) * -load,
                                    (self.nx, 2))
            self.BCs.append(ElAntiBB(tuple(self.boundingBoxIndices["top"].T), self.gridInfo, self.precisionPolicy, force))

    kw = {"lambda": lam, "mu": mu, "nx": nx, "ny": nx + 2, "nz": 0, "precision": precision,
          "lattice": m["D2Q8"](precision), "print_info_rate": 0}
    if lambda_field is not None:
        kw["lambda_field"], kw["mu_field"] = lambda_field, mu_field
    return El(**kw)
