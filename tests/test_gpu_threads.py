"""Two threads rendering two contexts at once.

The reference keeps one current context per thread (src/gl_api.c:14-23,
__thread) and shares nothing else between contexts; the host layer does the
same and each device context owns its stream and buffers.  Two threads, each
with its own context, render different scenes concurrently, several times
over; every frame must equal the oracle's frame of that scene (depth and
stencil bit-exact, colour within the per-frame bar)."""
import threading

import pytest

from conftest import MTGL_LIB
from mytinygl_amd import scenes as S
from mytinygl_amd.gl import GLLibrary
from test_gpu_parity import _check_frame

pytestmark = pytest.mark.gpu

JOBS = [
    [("textured_sphere_small", 160, 120, 1), ("stencil", 96, 72, 2), ("blend", 96, 72, 3)],
    [("overdraw_small", 160, 120, 2), ("lighting", 96, 72, 1), ("polygon_mode_mixed_dense", 96, 72, 2)],
]
REPS = 4


def test_two_threads_two_contexts(oracle):
    expected = {}
    for jobs in JOBS:
        for job in jobs:
            name, W, H, seed = job
            expected[job] = S.render(oracle, S.FEATURE_SCENES[name], W, H, seed)
    # one binding per thread: each keeps its own client arrays alive
    libs = [GLLibrary(MTGL_LIB, "mtgl") for _ in JOBS]
    start = threading.Barrier(len(JOBS))
    results, errors = {}, []

    def worker(k):
        try:
            g = libs[k]
            start.wait()
            for rep in range(REPS):
                for job in JOBS[k]:
                    name, W, H, seed = job
                    results[(k, rep, job)] = S.render(g, S.FEATURE_SCENES[name], W, H, seed)
        except Exception as e:   # reported by the main thread
            errors.append((k, repr(e)))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(len(JOBS))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in threads), "a rendering thread did not finish"
    assert not errors, errors
    assert len(results) == REPS * sum(len(j) for j in JOBS)
    for (k, rep, job), (planes, err) in sorted(results.items()):
        exp_planes, exp_err = expected[job]
        assert err == exp_err
        _check_frame(f"thread {k} rep {rep} {job[0]}", exp_planes, planes)
