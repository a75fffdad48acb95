"""Separate solve handles driven from concurrent host threads.

The reference's C ABI is single-threaded (one global FFTW planner,
src/fftw_helper.cpp:5-79). Here every SWESim handle owns its stream and device
buffers, and ctypes releases the GIL during each call, so handles used from
different Python threads run concurrently on the device. Their results must
equal those of the same solves run one after another, bit for bit (the same
kernels run on the same inputs; only their overlap differs).
"""
import os
import threading

import numpy as np
import pytest

from conftest import HALDANE, HMCASE0, make_param

pytestmark = pytest.mark.gpu


def _params(tmp):
    from hhgamd import workloads
    s4 = workloads.synthetic_file(tmp, 4)
    s8 = workloads.synthetic_file(tmp, 8)
    short = dict(tmax=0.02197 * 300 + 1e-9, intensity=5e11, pol=(0.6, 0.8, 0.0))
    return [
        make_param(HALDANE, nr=40),                                   # nb = 2 register pairs, FP64
        make_param(HMCASE0, nr=(24, 30), precision="fp32"),           # nb = 2, FP32
        make_param(s4, nr=24, precision="fp32", **short),             # row-separable, 4 bands
        make_param(s8, nr=(16, 20), **short),                         # row-separable, 8 bands, FP64
        make_param(HALDANE, nr=30, pol=(1.0, 0.0, 0.0), precision="fp32"),
        make_param(s4, nr=(20, 18), **short),
    ]


def _run(lib, prm):
    with lib.Solve(prm) as s:
        s.run()
        return s.current()[1]


def test_concurrent_handles_equal_sequential(tmp_path):
    from hhgamd import lib
    prms = _params(str(tmp_path))
    seq = [_run(lib, p) for p in prms]
    out = [None] * len(prms)
    errors = []

    def worker(i):
        try:
            out[i] = _run(lib, prms[i])
        except Exception as e:  # reported below with the case index
            errors.append((i, repr(e)))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(prms))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a solve thread did not finish"
    assert not errors, errors
    for i, (a, b) in enumerate(zip(seq, out)):
        print(f"case {i}: nt = {a.shape[0]}, max |diff| = {np.max(np.abs(a - b)):.3e}")
        assert np.array_equal(a, b), f"case {i} differs when run concurrently"

