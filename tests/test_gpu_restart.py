"""Checkpoint / restart (the reference's -c/-T options are stubs,
main.cpp:283-284): a run interrupted at step 10, written to disk and
restored into a fresh context continues bit for bit."""
import numpy as np
import pytest

from hydist_c_amd import abi, engine, scenarios as S
from parity_util import bits_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [abi.HD_FORM_FP64_PLAIN, abi.HD_FORM_FP64_FAST])
def test_restart_continues_bit_for_bit(tmp_path, form):
    print(f"[case] restart form={form}", flush=True)
    sc = S.tidal_flats(64, 40, seed=6)
    sc.sediment_start = sc.bed_change_start = 8.0
    a = engine.Simulation(sc, form=form, sediment=True, bed_change=True)
    a.init()
    a.step(10)
    ck = tmp_path / "state.hdck"
    a.save_checkpoint(ck)
    a.step(12)
    b = engine.Simulation(sc, form=form, sediment=True, bed_change=True)
    b.init()
    b.load_checkpoint(ck)
    assert b.t == 20.0
    b.step(12)
    assert a.t == b.t
    for f in ("U", "V", "Z", "FS", "DH", "KHOUOT"):
        assert bits_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()


def test_restore_rejects_other_mesh(tmp_path):
    a = engine.Simulation(S.open_box(40, 28, seed=3))
    a.init()
    ck = tmp_path / "a.hdck"
    a.save_checkpoint(ck)
    b = engine.Simulation(S.open_box(48, 28, seed=3))
    b.init()
    with pytest.raises(engine.HydistError, match="does not match"):
        b.load_checkpoint(ck)
    a.close()
    b.close()


def test_restore_rejects_a_damaged_checkpoint(tmp_path):
    """One flipped byte (in the header's time or anywhere in the payload),
    or a truncated file, is refused, and the context keeps its state."""
    print("[case] damaged checkpoint", flush=True)
    sc = S.open_box(40, 28, seed=3)
    a = engine.Simulation(sc)
    a.init()
    a.step(4)
    ck = tmp_path / "a.hdck"
    a.save_checkpoint(ck)
    raw = ck.read_bytes()
    b = engine.Simulation(sc)
    b.init()
    z0 = b.get("Z")
    for pos in (40, 200, len(raw) // 2, len(raw) - 1):  # 40: the time t
        bad = bytearray(raw)
        bad[pos] ^= 0x10
        (tmp_path / "bad.hdck").write_bytes(bytes(bad))
        with pytest.raises(engine.HydistError, match="checksum"):
            b.load_checkpoint(tmp_path / "bad.hdck")
    (tmp_path / "short.hdck").write_bytes(raw[:-8])
    with pytest.raises(engine.HydistError):
        b.load_checkpoint(tmp_path / "short.hdck")
    assert bits_equal(b.get("Z"), z0)
    b.load_checkpoint(ck)
    assert bits_equal(b.get("Z"), a.get("Z"))
    a.close()
    b.close()
