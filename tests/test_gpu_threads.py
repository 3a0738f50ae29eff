"""Separate lattice handles may be driven from separate host threads (each
handle has its own stream; the error message is thread-local).  Four lattices
stepped concurrently from four threads equal the same lattices stepped one
after another, value for value; one thread's argument error does not leak into
another thread's status."""
import threading

import numpy as np
import pytest

from chemsim_amd import Lattice, Params, capi
from chemsim_amd import workloads as W

pytestmark = pytest.mark.gpu


def make(ctx, k):
    if k % 2 == 0:  # launch-bound cavity: graph replay path
        n = 96 + 32 * k
        lat = Lattice(ctx, n, n, "f32", Params(tau=0.6, boundary=capi.CS_BB_HALFWAY, lid_uy=0.1),
                      W.cavity_flags(n, n))
        lat.init_equilibrium(*W.rest_state(n, n))
    else:  # periodic Taylor-Green: vector kernel
        n = 512 * k
        lat = Lattice(ctx, n, n, "f32" if k == 1 else "f16", Params(tau=0.8, edge=capi.CS_EDGE_PERIODIC))
        lat.init_equilibrium(*W.taylor_green(n, 0.05, np.float32))
    return lat


def test_concurrent_lattices_equal_sequential(ctx):
    steps = [300, 130, 257, 64]
    seq = []
    for k in range(4):
        with make(ctx, k) as lat:
            lat.step(steps[k])
            seq.append(lat.download_raw())
    lats = [make(ctx, k) for k in range(4)]
    errors = []

    def run(k):
        try:
            for _ in range(3):
                lats[k].step(steps[k] // 3)
            lats[k].step(steps[k] - 3 * (steps[k] // 3))
            lats[k].sync()
        except Exception as e:  # reported below
            errors.append((k, e))

    threads = [threading.Thread(target=run, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for k, lat in enumerate(lats):
        np.testing.assert_array_equal(lat.download_raw(), seq[k], err_msg=f"lattice {k}")
        lat.close()


def test_error_message_is_per_thread(ctx):
    msgs = {}

    def bad():
        try:
            capi.call("cs_set_cells_per_thread", 3)
        except capi.ChemsimError as e:
            msgs["bad"] = str(e)

    def good():
        capi.call("cs_set_cells_per_thread", 0)
        msgs["good"] = capi.last_error()

    t1 = threading.Thread(target=bad)
    t1.start()
    t1.join()
    t2 = threading.Thread(target=good)
    t2.start()
    t2.join()
    assert "cells per thread" in msgs["bad"]
    assert "cells per thread" not in msgs["good"]
