"""The converged-quantity fixtures (golden/converged.json) and the physics they must show.

The inviscid rotating sector is re-run on the oracle here (a few seconds) and must reproduce
its fixture; every fixture must show what the mathematics fixes for its case: mass
conservation between the inlet and outlet monitor planes (to the discretisation error of the
monitor's face interpolation on these coarse grids), no total-temperature change in the
passive bump channel and a total-pressure loss, work input (T0 rise) and compression in the
rotor, and an isentropic efficiency between 0 and 1."""
import json
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_converged as mc  # noqa: E402

with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "converged.json")) as _fh:
    GOLD = json.load(_fh)


def test_sector_euler_fixture_reproduces():
    q = mc.run_oracle("sector_euler", threads=min(8, os.cpu_count() or 1))
    g = GOLD["sector_euler"]
    assert q["iterations"] == g["iterations"]
    for k, v in g.items():
        if v is None:
            assert q[k] is None
        else:
            assert q[k] == pytest.approx(v, rel=1e-12), k


@pytest.mark.parametrize("name", ["sector_euler", "sector_rans", "bump65"])
def test_converged_mass_conservation(name):
    g = GOLD[name]
    tol = 5e-3 if name.startswith("sector") else 2e-3
    assert abs(g["mass_flow_out"] - g["mass_flow_in"]) <= tol * g["mass_flow_in"]
    assert g["mass_flow_in"] > 0


def test_reached_drops():
    assert GOLD["sector_euler"]["drop"] <= 1e-8
    assert GOLD["bump65"]["drop"] <= 1e-8
    assert GOLD["sector_rans"]["drop"] <= 1e-6
    assert GOLD["bump_2000"]["iterations"] == 2000


@pytest.mark.parametrize("name", ["bump65", "bump_2000"])
def test_bump_is_passive(name):
    g = GOLD[name]
    assert g["temperature_ratio"] == pytest.approx(1.0, abs=1e-3)
    assert 0.98 < g["pressure_ratio"] <= 1.0


@pytest.mark.parametrize("name", ["sector_euler", "sector_rans"])
def test_rotor_does_work(name):
    g = GOLD[name]
    assert g["temperature_ratio"] > 1.01
    assert g["pressure_ratio"] > 1.01
    assert 0.0 < g["efficiency"] < 1.0
