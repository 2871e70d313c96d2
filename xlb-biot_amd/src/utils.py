"""Host utilities used by the drivers of this path (reference ``src/utils.py``).

``Parameters`` (utils.py:17-76) converts between dimensional and lattice units for
the Terzaghi / 2-D consolidation set-ups.  ``downsample_field`` (utils.py:78-106)
is the identity at the default factor 1, which is what the solvers call; other
factors are an I/O feature that this path does not port.  ``save_fields_vtk``
(utils.py:144-196) writes the same files as the reference (cell data on a uniform
grid, one ``<prefix>_<timestep:07d>.vtk`` per call) as legacy binary VTK without
pyvista; ``save_error_history`` stores the error histories of the validation
drivers as ``.npz`` like ``examples/poroelasticity/periodic.py:243-259``.
"""
import functools
import os
from typing import Dict, NamedTuple

import numpy as np


class _Quantity(NamedTuple):
    value: float          # in the base units the quantity was given in (SI in the drivers)
    dims: Dict[str, int]  # base unit name -> integer power


class Parameters:
    """Physical parameters with their dimensions, and their values in a scaled unit
    system (reference ``src/utils.py:17-76``: same methods, arguments and errors).

    Base units (length, time, pressure, ...) are positive scale factors; a parameter
    carries a value and the integer powers of the base units it is measured in.  Its
    nondimensional value divides the value by each base unit raised to its power, one
    unit at a time in the order the powers were given (the drivers' lattice
    parameters come out of this, so the order of the divisions is kept)."""

    def __init__(self, units=None, parameters=None):
        self._scale: Dict[str, float] = {}
        self._quantities: Dict[str, _Quantity] = {}
        self.add_units(**dict(units or {}))
        self.add_parameters(**dict(parameters or {}))

    # ---- base units
    def set_unit(self, name, value):
        if not isinstance(value, float) or value <= 0.0:
            raise TypeError(f"Unit '{name}' must be a positive float.")
        self._scale[name] = value

    def add_units(self, **units):
        for name in units:
            self.set_unit(name, units[name])

    def get_unit(self, key):
        return self._scale[key]

    # ---- parameters
    def set_parameter(self, name, value, unit):
        if not isinstance(unit, dict):
            raise TypeError("unit must be a dict.")
        if not isinstance(value, float):
            raise TypeError("value must be a float.")
        # the first offending base unit, in the order the powers are given
        bad = next(((k, k not in self._scale) for k, p in unit.items()
                    if k not in self._scale or not isinstance(p, int)), None)
        if bad is not None:
            raise TypeError(f"unit['{bad[0]}'] must exist as base unit." if bad[1] else f"unit['{bad[0]}'] must be an int.")
        self._quantities[name] = _Quantity(value, unit)

    def add_parameters(self, **parameters):
        for name, spec in parameters.items():
            self.set_parameter(name, *spec)

    def get(self, key):
        return self._quantities[key].value

    def _scaled(self, q):
        return functools.reduce(lambda acc, d: acc / self._scale[d[0]] ** d[1], q.dims.items(), q.value)

    def get_nondimensional(self, key):
        return self._scaled(self._quantities[key])

    # ---- report
    @staticmethod
    def _dims_label(dims):
        return " ".join(k if p == 1 else f"{k}**{p}" for k, p in dims.items()) or "---"

    def print(self, nondimensional=True):
        rows = [("Parameter", "Unit", "Value"), None]
        rows += [(k, "< UNIT >", s) for k, s in self._scale.items()]
        rows += [(k, self._dims_label(q.dims), self._scaled(q) if nondimensional else q.value)
                 for k, q in self._quantities.items()]
        for row in rows:
            print("-" * 60 if row is None else f"{row[0]:>20} | {row[1]: ^20} | {row[2]}")


def downsample_field(field, factor, method="bicubic"):
    if factor == 1:
        return field
    raise NotImplementedError("down-sampling by a factor other than 1 is an I/O feature not ported to this path")


def _host(a):
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def save_fields_vtk(timestep, fields, output_dir=".", prefix="fields"):
    """Write ``fields`` (name -> ``(nx, ny)`` or ``(nx, ny, nz)`` array) as cell data of
    a uniform grid to ``<output_dir>/<prefix>_<timestep:07d>.vtk`` (legacy binary VTK,
    big-endian float64, x fastest as with ``flatten(order='F')``)."""
    items = [(k, _host(v)) for k, v in fields.items()]
    if not items:
        raise ValueError("no fields to save")
    dims = items[0][1].shape
    for k, v in items:
        if v.shape != dims:
            raise ValueError("All fields must have the same dimensions!")
    pts = tuple(d + 1 for d in dims) + ((1,) if len(dims) == 2 else ())
    ncell = int(np.prod(dims))
    path = os.path.join(output_dir, prefix + "_" + f"{timestep:07d}.vtk")
    with open(path, "wb") as fh:
        fh.write(b"# vtk DataFile Version 3.0\n")
        fh.write(f"{prefix} timestep {timestep}\n".encode())
        fh.write(b"BINARY\nDATASET STRUCTURED_POINTS\n")
        fh.write(("DIMENSIONS %d %d %d\n" % pts).encode())
        fh.write(b"ORIGIN 0 0 0\nSPACING 1 1 1\n")
        fh.write(f"CELL_DATA {ncell}\n".encode())
        for k, v in items:
            fh.write(f"SCALARS {k} double 1\nLOOKUP_TABLE default\n".encode())
            fh.write(np.asarray(v, dtype=">f8").flatten(order="F").tobytes())
            fh.write(b"\n")
    return path


def read_fields_vtk(path):
    """Read back a file written by :func:`save_fields_vtk`: ``(dims, {name: array})``."""
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0
    fields, dims = {}, None
    ncell = None

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        text = data[pos:end].decode()
        pos = end + 1
        return text
    for _ in range(4):
        line()
    dims = tuple(int(t) - 1 for t in line().split()[1:])
    line(), line()
    ncell = int(line().split()[1])
    shape = tuple(d for d in dims if d > 0) if dims[-1] == 0 else dims
    while pos < len(data):
        hdr = line()
        if not hdr.startswith("SCALARS"):
            continue
        name = hdr.split()[1]
        line()
        arr = np.frombuffer(data[pos:pos + 8 * ncell], dtype=">f8")
        pos += 8 * ncell + 1
        fields[name] = arr.reshape(shape, order="F")
    return shape, fields


def save_error_history(filename, **arrays):
    """Error histories / relative errors of a validation run, ``np.savez_compressed``
    with the driver's key names (e.g. ``err_l2_eta``, ``rel_err_l2_p``)."""
    np.savez_compressed(filename, **{k: _host(v) for k, v in arrays.items()})
