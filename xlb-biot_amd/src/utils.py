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
import os

import numpy as np


class Parameters:
    def __init__(self, units=None, parameters=None):
        self._units = {}
        self._params = {}
        self.add_units(**(units or {}))
        self.add_parameters(**(parameters or {}))

    def set_unit(self, name, value):
        if not isinstance(value, float) or value <= 0.0:
            raise TypeError(f"Unit '{name}' must be a positive float.")
        self._units[name] = value

    def add_units(self, **kwargs):
        for k, v in kwargs.items():
            self.set_unit(k, v)

    def set_parameter(self, name, value, unit):
        if not isinstance(unit, dict):
            raise TypeError("unit must be a dict.")
        if not isinstance(value, float):
            raise TypeError("value must be a float.")
        for k, v in unit.items():
            if k not in self._units:
                raise TypeError(f"unit['{k}'] must exist as base unit.")
            if not isinstance(v, int):
                raise TypeError(f"unit['{k}'] must be an int.")
        self._params[name] = (value, unit)

    def add_parameters(self, **kwargs):
        for k, v in kwargs.items():
            self.set_parameter(k, *v)

    def get_unit(self, key):
        return self._units[key]

    def get(self, key):
        return self._params[key][0]

    def get_nondimensional(self, key):
        value, unit = self._params[key]
        for k, v in unit.items():
            value /= self._units[k] ** v
        return value

    def print(self, nondimensional=True):
        print(f"{'Parameter':>20} | {'Unit': ^20} | {'Value'}")
        print('-' * 60)
        for k, v in self._units.items():
            print(f"{k:>20} | {'< UNIT >': ^20} | {v}")
        for k, (v, u) in self._params.items():
            unit = ""
            for uk, uv in u.items():
                if nondimensional:
                    v /= self._units[uk] ** uv
                unit += (" " if unit else "") + uk + (f"**{uv}" if uv != 1 else "")
            print(f"{k:>20} | {unit or '---': ^20} | {v}")


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
