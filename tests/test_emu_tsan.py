"""The library's threads under ThreadSanitizer, on the CPU: the
emulated build (tests/emu) compiled with -fsanitize=thread runs a server
with whole-step graphs while three client threads submit, poll, read
errors, cancel (queued and running runs), wait, fetch results and release,
so the worker, the context-making helper and the clients' API calls all
meet.  A data race or lock-order problem in the host code ends the child
with ThreadSanitizer's report.  (The emulator announces its lane fibers to
ThreadSanitizer, and its launches and allocations take one lock, as on
one queue.  The clients wait without a time limit: the limited wait goes
through pthread_cond_clockwait, which this compiler's ThreadSanitizer
does not intercept, so it would report the mutex as locked twice; the
subprocess's own time limit bounds the run instead.)"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

SERVER = r"""
import sys, threading, time
sys.path.insert(0, {root!r})
from hydist_c_amd import abi, engine, scenarios as S
engine.LIB_PATH = {lib!r}
engine._lib = None
FAST, PLAIN = abi.HD_FORM_FP64_FAST, abi.HD_FORM_FP64_PLAIN
big = S.open_box(36, 24, seed=5)
srv = engine.Server(hbm_budget=engine.footprint(big) + 1, max_active=8, steps_per_quantum=2, use_graph=True)
first = srv.submit(big, 20.0, form=PLAIN)  # still running when the clients submit
errors = []

def client(k):
    try:
        ids = [srv.submit(S.open_box(20, 16, seed=20 + 2 * k + j), 6.0, form=FAST, sediment=True,
                          bed_change=True) for j in range(2)]
        queued = srv.submit(S.open_box(24, 18, seed=90 + k), 400.0)
        srv.cancel(queued)
        for r in ids:
            while srv.poll(r)[0] < engine.RUN_DONE:
                srv.error(r)
                time.sleep(0.01)
            assert srv.wait(r) == engine.RUN_DONE, (k, r, srv.error(r))
            srv.result(r, "Z")
            srv.release(r)
        assert srv.wait(queued) == engine.RUN_CANCELLED
        if k == 0:  # a run cancelled while it is on the GPU
            r = srv.submit(S.open_box(20, 16, seed=7), 4000.0)
            while srv.poll(r)[0] != engine.RUN_RUNNING:
                time.sleep(0.01)
            srv.cancel(r)
            assert srv.wait(r) == engine.RUN_CANCELLED
    except Exception as e:
        errors.append(repr(e))

ts = [threading.Thread(target=client, args=(k,)) for k in range(3)]
for t in ts:
    t.start()
for t in ts:
    t.join()
assert srv.wait(first) == engine.RUN_DONE
st = srv.stats()
srv.close()
assert not errors, errors
assert st["completed"] == 7 and st["cancelled"] == 4 and st["failed"] == 0 and st["batched"] >= 2, st
print("server ok", st, flush=True)
"""


CONTEXTS = r"""
import sys, threading
sys.path.insert(0, {root!r})
from hydist_c_amd import abi, engine, scenarios as S
engine.LIB_PATH = {lib!r}
engine._lib = None
errors = []

def own(k, form):
    try:
        sc = S.open_box(16, 12, seed=40 + k)
        sc.sediment_start = sc.bed_change_start = 2.0
        g = engine.Simulation(sc, form=form, sediment=True, bed_change=True)
        g.init()
        if k == 0:
            engine.autotune(g)
        g.step(2)
        g.get("Z")
        g.close()
    except Exception as e:
        errors.append(repr(e))

forms = [abi.HD_FORM_FP64_FAST, abi.HD_FORM_FP64_PLAIN, abi.HD_FORM_FP32_REFINE]
ts = [threading.Thread(target=own, args=(k, f)) for k, f in enumerate(forms)]
for t in ts:
    t.start()
for t in ts:
    t.join()
assert not errors, errors
print("contexts ok", flush=True)
"""


@pytest.fixture(scope="module")
def tsan_build():
    sys.path.insert(0, os.path.join(HERE, "emu"))
    import build_emu
    lib = build_emu.build(sanitize="thread")
    rt = subprocess.run(["g++", "-print-file-name=libtsan.so"], capture_output=True, text=True).stdout.strip()
    assert os.path.isabs(rt) and os.path.exists(rt), rt
    return lib, rt


def test_serving_loop_under_tsan(tsan_build):
    lib, rt = tsan_build
    env = dict(os.environ, LD_PRELOAD=rt, TSAN_OPTIONS="halt_on_error=1:exitcode=66")
    r = subprocess.run([sys.executable, "-c", SERVER.format(root=ROOT, lib=lib)], capture_output=True, text=True,
                       timeout=900, env=env)
    assert r.returncode == 0 and "server ok" in r.stdout, r.stdout[-2000:] + r.stderr[-6000:]
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-6000:]


def test_contexts_on_threads_under_tsan(tsan_build):
    """Contexts of the three forms created, tuned, stepped and destroyed on
    three threads at once (a context is used by one thread at a time;
    different contexts share no host state)."""
    lib, rt = tsan_build
    env = dict(os.environ, LD_PRELOAD=rt, TSAN_OPTIONS="halt_on_error=1:exitcode=66")
    r = subprocess.run([sys.executable, "-c", CONTEXTS.format(root=ROOT, lib=lib)], capture_output=True, text=True,
                       timeout=900, env=env)
    assert r.returncode == 0 and "contexts ok" in r.stdout, r.stdout[-2000:] + r.stderr[-6000:]
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-6000:]
