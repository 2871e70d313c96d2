"""Host-side parts of the study drivers (``xlb-biot_amd/examples``): the mode sums,
the Terzaghi non-dimensionalisation, file names and the sweeps' study lists,
against the test studies and the run scripts' loops."""
import os
import sys

import numpy as np
import pytest

import orders as O
import problems as P

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "xlb-biot_amd",
                                "examples"))
import common  # noqa: E402
import periodic  # noqa: E402
import sweeps  # noqa: E402
import terzaghi  # noqa: E402


@pytest.mark.parametrize("kls", [[[], [2, 2], [], []], [[], [], [], [2, 2]], [[2, 2, 4, 6], [], [1, 3], []]])
def test_mode_sum_matches_the_sum_function(kls):
    np.testing.assert_allclose(common.mode_sum(12, 10, kls, lambda k, l: 0.5 + k * l),
                               P.sum_function(12, 10, kls, lambda k, l: 0.5 + k * l), rtol=0, atol=1e-13)


@pytest.mark.parametrize("n", [8, 32])
def test_terzaghi_nondimensional_parameters(n):
    # Parameters (utils.py:17-76) reproduce the hand-derived lattice parameters
    prm = terzaghi.parameters(1.0)
    prm.add_units(L=prm.get("L_tot"), H=prm.get("eta0"), T=prm.get("T_tot"), P=prm.get("p0"))
    prm.add_units(L=prm.get("L_tot") / n, H=prm.get("eta0"), T=prm.get("T_tot") * 0.25 / n ** 2, P=prm.get("p0"))
    s = O.terzaghi_setup(n)
    got = {"lam": prm.get_nondimensional("lambda"), "mu": prm.get_nondimensional("mu"),
           "kappa": prm.get_nondimensional("kappa"), "force": prm.get_nondimensional("force"),
           "eta_inf": prm.get_nondimensional("etaInf"), "epsilon": prm.get_unit("L") / prm.get_unit("H"),
           "nt": int(prm.get_nondimensional("T_tot"))}
    for k, v in got.items():
        assert v == pytest.approx(s[k], rel=1e-12), k


def test_result_file_names():
    assert os.path.basename(periodic.npz_name(".", 0.5, 20, 20, 1, "centered", False)) == \
        "error_E0.110_nu0.800_K0.100_a0.500_nx020_ne020_l001_centered.npz"
    assert os.path.basename(terzaghi.npz_name(".", 0.9, 64, 3, 6, "implicit", True)) == \
        "error_a0.900_nx064_ne003_l006_implicit_altgrad.npz"


def test_depth_and_refinement_rules():
    assert common.refinement_list(4, 5, -1) == [4, 8, 16, 32, 64, 128]
    assert common.refinement_list(10, 2, 1) == [10, 20, 30]
    assert common.depth_and_steps(128, -1, 3) == (7, 3)
    assert common.depth_and_steps(24, -1, 2) == (4, 2)
    assert common.depth_and_steps(30, 1, 0.5) == (1, 15)


def test_sweep_plans_follow_the_run_scripts():
    assert len(sweeps.plan("periodic")) == 45 and len(sweeps.plan("periodic", coarse=True)) == 90
    assert len(sweeps.plan("periodic_mg")) == 114          # explicit only for alpha < 0.8, implicit < 0.9
    assert len(sweeps.plan("periodic_mg_altgrad")) == 132  # explicit only for alpha < 0.9
    assert len(sweeps.plan("terzaghi")) == 54 and len(sweeps.plan("terzaghi", coarse=True)) == 108
    mod, kw = sweeps.plan("periodic_mg_altgrad")[0]
    assert mod is periodic and kw["gradient"] and kw["levels"] == -1 and kw["refinements"] == 5
