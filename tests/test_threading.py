"""Concurrent calls on one simulator equal sequential ones
(src/qarpx/python/bindings.cpp:1364-1508 releases the GIL around every call;
the reference's tests/test_threading.py:360-390 checks the same).  Every
library entry point takes one process-wide lock and ctypes releases the GIL
around each call, so Python threads interleave whole calls; error messages
are per thread."""
import threading

import numpy as np
import pytest

from openqarp_amd import workloads as W
from openqarp_amd.ir import Command, GateType
from test_gpu_parity import random_circuit


@pytest.mark.gpu
def test_concurrent_calls_equal_sequential():
    from openqarp_amd.simulator import QarpSimulator
    sim = QarpSimulator()
    jobs = []
    for i in range(4):
        n = 12 + i
        cmds = W.brickwork_haar(n, 4, i) + random_circuit(n, 20, i)
        obs = [(t, float(np.real(c))) for t, c in W.random_pauli_sum(n, 12, i)]
        jobs.append(lambda c=cmds, n=n: sim.statevector(c, n))
        jobs.append(lambda c=cmds, n=n, o=obs: sim.expectation(c, n, o))
        jobs.append(lambda c=cmds, n=n: sim.run(c + [Command(GateType.Measure, q, cbits=[q]) for q in range(n)],
                                                n, 500, seed=11).counts)
    bad = [Command(GateType.X, 40)]
    seq = [f() for f in jobs]
    out = [None] * len(jobs)
    errors = []

    def work(j, rounds):
        for _ in range(rounds):
            r = jobs[j]()
            if out[j] is None:
                out[j] = r

    def fail_loop():
        for _ in range(20):
            try:
                sim.statevector(bad, 4)
            except ValueError as e:
                errors.append(str(e))

    threads = [threading.Thread(target=work, args=(j, 3)) for j in range(len(jobs))]
    threads.append(threading.Thread(target=fail_loop))
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for a, b in zip(seq, out):
        if isinstance(a, np.ndarray):
            assert np.array_equal(a, b)
        else:
            assert a == b
    assert len(errors) == 20 and len(set(errors)) == 1
