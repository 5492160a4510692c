"""The drop-in path's environment switches are parsed by name: an unknown
HYDIST_FORM or HYDIST_Q11 ends the compat driver with status 2 and a
message before any device work (so a misspelt form never runs as
another).  No GPU needed."""
import os
import subprocess

import pytest

from hydist_c_amd import inputs, scenarios as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "_build", "compat_driver")


@pytest.mark.parametrize("var,value", [("HYDIST_FORM", "fastest"), ("HYDIST_FORM", "3"),
                                       ("HYDIST_Q11", "yes")])
def test_unknown_switch_value_stops_before_device_work(tmp_path, var, value):
    if not os.path.exists(DRIVER):
        pytest.skip("compat driver not built")
    d = inputs.write_scenario_inputs(S.open_box(20, 16), str(tmp_path / "Inputs"))
    env = dict(os.environ)
    env[var] = value
    r = subprocess.run([DRIVER, "-f", d, "-h", "0", "-m", "0", "-s", "4"], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=60, env=env)
    assert r.returncode == 2, (r.returncode, r.stdout[-500:])
    assert f"{var}={value}" in r.stdout
