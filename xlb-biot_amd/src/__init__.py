"""MI355X (gfx950) lattice-Boltzmann solvers for linear elasticity (D2Q8), Darcy flow
(D2Q9) and their Biot poroelastic coupling, with the elasticity multigrid V-cycle.

Drop-in package for the reference's ``src`` package: ``from src.elasticity import
LBMElasticity`` etc. with ``xlb-biot_amd`` on ``sys.path``.  All stepping runs in
the HIP kernels of ``lib/libxlb_biot.so`` (C ABI in ``include/xlb_biot.h``).
"""
