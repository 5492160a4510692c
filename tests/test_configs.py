"""The benchmark configuration files (hydist_c_amd/configs/c1-c5.json)
describe the same inputs as the generators the tests and the bench use:
each file's generator, arguments, coefficient overrides and options give
bit for bit the scenario of config1() / config2() / ensemble_member(k) /
serving_request(k), and each file carries the run parameters of its
configuration."""
import numpy as np
import pytest

from hydist_c_amd import scenarios as S


def same_scenario(a, b):
    assert (a.M, a.N, a.total_time, tuple(a.boundary_type)) == (b.M, b.N, b.total_time, tuple(b.boundary_type))
    for f in ("h", "hsnham", "visc", "Fw"):
        x, y = getattr(a, f), getattr(b, f)
        assert (x is None and y is None) or np.array_equal(x, y), f
    for k in range(4):
        assert np.array_equal(np.asarray(a.bc[k]), np.asarray(b.bc[k]))
    assert (a.CC is None) == (b.CC is None)
    for k in range(4 if a.CC is not None else 0):
        assert np.array_equal(np.asarray(a.CC[k]), np.asarray(b.CC[k]))
    assert a.coeff_overrides == b.coeff_overrides
    assert (a.sediment_start, a.bed_change_start) == (b.sediment_start, b.bed_change_start)


@pytest.mark.parametrize("name,gen", [("c1", S.config1), ("c2", S.config2)])
def test_config_files_give_the_generators_scenarios(name, gen):
    cfg = S.load_config(name)
    same_scenario(S.config_scenario(cfg), gen())
    assert cfg["options"]["Tmax"] == 3600.0


def test_ensemble_config_members():
    cfg = S.load_config("c4")
    assert cfg["members"] == 256 and set(cfg["forms"]) == {"fast", "fp32"}
    for k in (0, 255):
        same_scenario(S.config_scenario(cfg, k), S.ensemble_member(k))


def test_serving_config_requests():
    cfg = S.load_config("c5")
    assert cfg["arrival_rate_per_s"] == 20.0 and cfg["window_s"] == 600.0
    for k in range(4):
        a, b = S.config_scenario(cfg, k), S.serving_request(k, fp64_form=1)
        same_scenario(a[0], b[0])
        assert a[1:] == b[1:]


def test_braided_config_parameters():
    cfg = S.load_config("c3")
    assert cfg["generator"] == "config3" and cfg["args"] == {"seed": 5, "sediment": True}
    assert cfg["coeffs"]["dT"] == 1.0 and cfg["options"]["Tmax"] == 600.0
