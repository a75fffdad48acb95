"""The allocation-free run, SWESim_run_simulation_static, called through the
C ABI: its J(t) equals SWESim_run_simulation's on a twin handle, run after
run, for a 2-band FP64 solve and then a smaller 4-band FP32 one; a solve whose
current partials exceed the static scratch is refused with an error."""
import ctypes as ct

import numpy as np
import pytest

from conftest import HALDANE, make_param

pytestmark = pytest.mark.gpu


def run_static(s):
    f = s.lib.SWESim_run_simulation_static
    f.argtypes = [ct.c_void_p, np.ctypeslib.ndpointer(dtype=np.float64, flags="C")]
    f.restype = ct.c_int
    out = np.zeros(6 * s.nt)
    rc = f(s.ptr, out)
    err = s.lib.SWESim_get_error(s.ptr).decode()
    return rc, out.reshape(s.nt, 6).view(np.complex128), err


def test_static_runs_equal_run_simulation(tmp_path):
    from hhgamd import lib, tbgen
    syn4 = str(tmp_path / "syn4_tb.dat")
    tbgen.write_tb(syn4, *tbgen.synthetic(4, seed=61))
    cases = [make_param(HALDANE, nr=(120, 120)),  # 4096 steps x 225 wave partials x 3 components
             dict(make_param(syn4, nr=(12, 16), tmax=10.0, tmax_field=8.0, pol=(0.6, 0.8, 0.0)), precision="fp32")]
    for p in cases:
        with lib.Solve(p) as a, lib.Solve(p) as b:
            for run in range(2):  # the second run continues from the state of the first
                rc, Js, err = run_static(a)
                assert rc == 0, err
                b.run()
                Jr = b.current_run(run)[1]
                print(f"{p['nr1']}x{p['nr2']} nb={a.nb} run {run}: static vs run_simulation max diff "
                      f"{np.abs(Js - Jr).max():.1e}")
                assert np.array_equal(Js, Jr)
            assert np.array_equal(a.rho_k(), b.rho_k())


def test_static_run_refuses_partials_beyond_the_scratch():
    from hhgamd import lib
    with lib.Solve(make_param(HALDANE, nr=(240, 240))) as s:  # 900 wave partials x 4096 steps x 3: 88 MB
        rc, _, err = run_static(s)
        assert rc != 0 and "static scratch" in err
        s.run()  # the allocating entry point still runs it
        assert np.isfinite(s.current()[1]).all()
