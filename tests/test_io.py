"""I/O formats of the drivers: the VTK field writer and the ``.npz`` error histories."""
import numpy as np

from src.utils import Parameters, read_fields_vtk, save_error_history, save_fields_vtk


def test_vtk_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    eta_x = rng.standard_normal((7, 5))
    p = rng.standard_normal((7, 5))
    path = save_fields_vtk(12, {"eta_x": eta_x, "p": p}, str(tmp_path))
    assert path.endswith("fields_0000012.vtk")
    head = open(path, "rb").read(200)
    assert b"DIMENSIONS 8 6 1" in head and b"CELL_DATA 35" in head
    shape, fields = read_fields_vtk(path)
    assert shape == (7, 5)
    np.testing.assert_array_equal(fields["eta_x"], eta_x)
    np.testing.assert_array_equal(fields["p"], p)


def test_error_history_npz(tmp_path):
    f = tmp_path / "error_E0.110_nu0.800.npz"
    save_error_history(str(f), err_l2_eta=np.arange(3.0), rel_err_l2_p=np.float64(0.25))
    d = np.load(str(f))
    assert d["err_l2_eta"].tolist() == [0.0, 1.0, 2.0] and float(d["rel_err_l2_p"]) == 0.25


def test_parameters_errors_and_report(capsys):
    # the validation messages of utils.py:24-43 and the table of utils.py:61-76
    import pytest
    P = Parameters(units={"L": 2.0})
    for bad in (1, -1.0, 0.0):
        with pytest.raises(TypeError, match="Unit 'T' must be a positive float."):
            P.set_unit("T", bad)
    with pytest.raises(TypeError, match="unit must be a dict."):
        P.set_parameter("a", 1.0, [("L", 1)])
    with pytest.raises(TypeError, match="value must be a float."):
        P.set_parameter("a", 1, {"L": 1})
    with pytest.raises(TypeError, match=r"unit\['T'\] must exist as base unit."):
        P.set_parameter("a", 1.0, {"L": 1, "T": 1})
    with pytest.raises(TypeError, match=r"unit\['L'\] must be an int."):
        P.set_parameter("a", 1.0, {"L": 1.0})
    P.add_units(T=4.0)
    P.add_parameters(v=(8.0, {"L": 1, "T": -1}), n=(3.0, {}), a=(16.0, {"L": 2}))
    assert P.get("v") == 8.0 and P.get_unit("T") == 4.0
    assert P.get_nondimensional("v") == 16.0 and P.get_nondimensional("a") == 4.0 and P.get_nondimensional("n") == 3.0
    P.print()
    P.print(nondimensional=False)
    lines = capsys.readouterr().out.splitlines()
    head = [f"{'Parameter':>20} | {'Unit': ^20} | Value", "-" * 60,
            f"{'L':>20} | {'< UNIT >': ^20} | 2.0", f"{'T':>20} | {'< UNIT >': ^20} | 4.0"]
    assert lines[:7] == head + [f"{'v':>20} | {'L T**-1': ^20} | 16.0", f"{'n':>20} | {'---': ^20} | 3.0",
                                f"{'a':>20} | {'L**2': ^20} | 4.0"]
    assert lines[7:] == head + [f"{'v':>20} | {'L T**-1': ^20} | 8.0", f"{'n':>20} | {'---': ^20} | 3.0",
                                f"{'a':>20} | {'L**2': ^20} | 16.0"]


def test_parameters_nondimensional():
    # utils.py:17-76: value / prod(unit^power)
    P = Parameters()
    P.add_units(L=2.0, T=4.0, P=10.0)
    P.add_parameters(kappa=(1e-9, {"L": 2, "T": -1, "P": -1}), E=(1e5, {"P": 1}))
    np.testing.assert_allclose(P.get_nondimensional("kappa"), 1e-9 / (2.0 ** 2 * 4.0 ** -1 * 10.0 ** -1))
    np.testing.assert_allclose(P.get_nondimensional("E"), 1e4)
