"""Biot coupler of the elasticity and Darcy oracles.  Oracle / test use only.

Restates ``src/poroelasticity.py:44-339`` (construction, ``run`` initialisation,
the coupled ``step``) on torch-CPU arrays.
"""
import torch
from .bc import gathered


class OraclePoroelasticity:
    def __init__(self, ElasticityCls, DarcyCls, **kwargs):
        # poroelasticity.py:44-70
        kw_el = dict(kwargs)
        kw_el["lattice"] = kwargs["lattice_elasticity"]
        kw_el["MG_levels"] = kwargs.get("elasticity_MG_levels", 1)
        self.elasticity = ElasticityCls(**kw_el)
        kw_d = dict(kwargs)
        kw_d["lattice"] = kwargs["lattice_darcy"]
        self.darcy = DarcyCls(**kw_d)
        self.elasticitySteps = kwargs.get("elasticity_steps", 1)
        self.alpha = kwargs.get("alpha")
        self.epsilon = kwargs.get("epsilon", 1 / self.elasticity.nx)
        method = kwargs.get("coupling_method", "centered")
        self.couplingRatio = {"explicit": 0.0, "implicit": 1.0, "centered": 0.5}[method]
        self.alternative_gradient = kwargs.get("alternative_gradient", False)
        self.ioRate = kwargs.get("io_rate", 0)
        self.useDynamicForce = self.get_elastic_force(0) is not None
        self.useDynamicSource = self.get_darcy_source(0) is not None

    def get_elastic_force(self, timestep):
        return None

    def get_darcy_source(self, timestep):
        return None

    def initialize_macroscopic_fields(self):
        return None

    def output_data(self, **kwargs):
        pass

    def initial_state(self):
        """poroelasticity.py:229-252."""
        el, da = self.elasticity, self.darcy
        dt_div_eta = self.initialize_macroscopic_fields()
        if dt_div_eta is None:
            dt_div_eta = torch.zeros((da.nx, da.ny, 1), dtype=da.precisionPolicy.compute_dtype)
        f_darcy = da.assign_fields_sharded()
        source = (-self.alpha / self.epsilon) * dt_div_eta
        if self.useDynamicSource:
            source = source + torch.as_tensor(self.get_darcy_source(0), dtype=source.dtype)
        f_darcy = f_darcy - 0.5 * source * torch.as_tensor(da.w, dtype=source.dtype)
        pressure = da.update_macroscopic(f_darcy, source)
        f_el = el.assign_fields_sharded()
        if self.alternative_gradient:
            force = da.gradient(pressure, -self.alpha * self.epsilon)
        else:
            _, pressure_grad = da.update_macroscopic(f_darcy, gradient=True)
            pressure_grad = pressure_grad + da.gradient(source, 0.5)
            force = (-self.alpha * self.epsilon) * pressure_grad
        if self.useDynamicForce:
            force = force + torch.as_tensor(self.get_elastic_force(0), dtype=force.dtype)
        eta, sigma = el.initialize_macroscopic_fields()
        if eta is not None and sigma is not None:
            f_el = el.equilibrium(eta, sigma, external_force=force)
        return f_el, f_darcy, dt_div_eta

    def step(self, f_elasticity, f_darcy, dt_div_eta, timestep):
        # poroelasticity.py:303-339
        el, da = self.elasticity, self.darcy
        th, a, eps = self.couplingRatio, self.alpha, self.epsilon
        fprev = f_elasticity.clone()
        source_base = ((1.0 - th) * -a / eps) * dt_div_eta
        if self.useDynamicSource:
            source_base = source_base + torch.as_tensor(self.get_darcy_source(timestep), dtype=source_base.dtype)
        if self.alternative_gradient:
            pressure_grad_base = da.gradient(da.update_macroscopic(f_darcy, source_base))
        else:
            _, pressure_grad_base = da.update_macroscopic(f_darcy, gradient=True)
            pressure_grad_base = pressure_grad_base + da.gradient(source_base, 0.5)
        force = (-a * eps) * pressure_grad_base - da.gradient(el.update_divergence(fprev), 0.5 * th * a ** 2)
        if self.useDynamicForce:
            force = force + torch.as_tensor(self.get_elastic_force(timestep), dtype=force.dtype)

        def nonlinear_force(f_el):
            return da.gradient(el.update_divergence(f_el), 0.5 * th * a ** 2)

        for _ in range(self.elasticitySteps):
            f_elasticity = el.step(f_elasticity, timestep, False, force, nonlinear_force=nonlinear_force)[0]
        eta, sigma = el.update_macroscopic(f_elasticity, force + nonlinear_force(f_elasticity))
        dt_div_eta = el.update_divergence(f_elasticity - fprev)
        source = source_base - (th * a / eps) * dt_div_eta
        pressure = da.update_macroscopic(f_darcy, source)
        if self.alternative_gradient:
            pressure_grad = da.gradient(pressure)
        else:
            pressure_grad = pressure_grad_base - da.gradient(dt_div_eta, 0.5 * th * a / eps)
        f_darcy = da.step(f_darcy, timestep, False, source)[0]
        return f_elasticity, f_darcy, eta, sigma, dt_div_eta, pressure, pressure_grad

    def run(self, t_max):
        # poroelasticity.py:208-301: t_max + 1 physical steps
        f_el, f_d, dt_div_eta = self.initial_state()
        for timestep in range(0, t_max + 1):
            f_el, f_d, eta, sigma, dt_div_eta, pressure, pressure_grad = self.step(f_el, f_d, dt_div_eta, timestep)
            if self.ioRate > 0 and (timestep % self.ioRate == 0 or timestep == t_max):
                self.output_data(timestep=timestep, eta=gathered(eta), sigma=gathered(sigma),
                                 dt_div_eta=gathered(dt_div_eta), pressure=gathered(pressure),
                                 pressure_grad=gathered(pressure_grad),
                                 f_poststreaming_elasticity=f_el, f_poststreaming_darcy=f_d)
        return f_el, f_d
