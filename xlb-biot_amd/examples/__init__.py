"""Drivers of the reference's poroelastic studies on this package's solvers.

``periodic`` (examples/poroelasticity/periodic.py) and ``terzaghi``
(examples/poroelasticity/Terzaghi.py) keep the drivers' command lines, result
folders, ``.npz`` file names and keys, so their error histories can be compared
file for file; ``sweeps`` runs the parameter grids of the run scripts
(runPeriodic.sh, runPeriodicMG.sh, runPeriodicMG_altgrad.sh, runTerzaghi.sh) in
one process on one GPU.
"""
