"""hd_gtsv_long_noalloc, the entry point that allocates no device memory,
against hd_gtsv_long (the same kernels with work memory allocated per
call): bit for bit, on two calls in a row, the second with a smaller
system (so the fixed work array keeps the first call's contents when the
second starts), and sizes past the array's limit rejected."""
import numpy as np
import pytest

from hydist_c_amd import engine
from test_gpu_gtsv import EMU, Dev, make_systems
from parity_util import bits_equal

pytestmark = pytest.mark.gpu

SIZES = [12289, 5000] if EMU else [engine.GTSV_NOALLOC_MAX_M, 100_003]


def solve(fn, sysv, m, precision):
    dl, d, du, b = sysv
    bufs = [Dev(a[0]) for a in (dl, d, du, b)]
    xb = Dev(np.full(m + 5, 7.0))
    fn(*[q.ptr for q in bufs], xb.ptr, m, precision)
    x = xb.get()
    for q in bufs + [xb]:
        q.free()
    return x


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32r"])
def test_noalloc_equals_allocating_entry_point(precision):
    for k, m in enumerate(SIZES):
        print(f"[case] gtsv_long_noalloc call {k + 1}: m={m} precision={precision}", flush=True)
        sysv = make_systems(np.random.default_rng(300 + k), m, 1, m + 5)
        a = solve(engine.gtsv_long_noalloc, sysv, m, precision)
        b = solve(engine.gtsv_long, sysv, m, precision)
        assert np.all(a[m:] == 7.0), "wrote past m"
        assert bits_equal(a, b), (m, float(np.max(np.abs(a - b))))


def test_noalloc_rejects_more_rows_than_its_array():
    print("[case] gtsv_long_noalloc limit", flush=True)
    one = Dev(np.zeros(8))
    p = one.ptr
    with pytest.raises(engine.HydistError, match="exceeds"):
        engine.gtsv_long_noalloc(p, p, p, p, p, engine.GTSV_NOALLOC_MAX_M + 1, 0)
    one.free()
