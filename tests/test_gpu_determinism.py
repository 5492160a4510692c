"""Run-twice check: the same scenario run twice, in two contexts created
one after the other and stepped interleaved, gives the same bits in every
form, for every state field (the kernels have no races: every cell and
every line is owned by one thread or one workgroup, and no result depends
on the order in which workgroups finish)."""
import pytest

from hydist_c_amd import abi, engine, scenarios as S
from parity_util import bits_equal

pytestmark = pytest.mark.gpu

FORMS = {abi.HD_FORM_FP64_PLAIN: "plain", abi.HD_FORM_FP64_FAST: "fp64_fast", abi.HD_FORM_FP32_REFINE: "fp32_refine"}
FIELDS = ["U", "V", "Z", "HMOI", "FS", "DH", "KHOUOT", "VTH", "KX", "KY", "QBX", "QBY"]


@pytest.mark.parametrize("form", list(FORMS), ids=FORMS.get)
def test_two_runs_give_the_same_bits(form):
    sc = S.meander_islands(96, 64, n_islands=4)
    sc.sediment_start = sc.bed_change_start = 0.0
    print(f"[case] run twice {FORMS[form]} {sc.name}", flush=True)
    a = engine.Simulation(sc, form=form, sediment=True, bed_change=True)
    b = engine.Simulation(sc, form=form, sediment=True, bed_change=True)
    a.init()
    b.init()
    for _ in range(8):
        a.step()
        b.step()
    for f in FIELDS:
        assert bits_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()
