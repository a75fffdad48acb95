"""libwannier.so without a GPU: exported ABI, host setup against the oracle,
error behaviour, and the pySWE mirror's argument handling."""
import ctypes as ct
import os
import subprocess

import numpy as np
import pytest

from conftest import HALDANE, HMCASE0, PKG, make_param

REFERENCE_ABI = [  # src/cwrapper.cpp:7-27
    "SWESim_new", "SWESim_run_simulation", "SWESim_test_files", "SWESim_set_path_tb", "SWESim_set_settings",
    "SWESim_init", "SWESim_restart", "SWESim_get_current", "SWESim_delete", "Settings_new", "Settings_set_nr1",
    "Settings_set_nr2", "Settings_set_nr3", "Settings_set_tmax", "Settings_set_dt", "Settings_set_intensity",
    "Settings_set_lambda", "Settings_set_tmax_field", "Settings_set_pol_vec", "Settings_set_phi_vec",
]


@pytest.fixture(scope="module")
def native():
    from hhgamd import lib
    if not os.path.exists(lib.LIB_PATH):
        lib.build()
    return lib


def test_header_declares_reference_abi(native):
    declared = native.declared_symbols()
    for name in REFERENCE_ABI:
        assert name in declared


def test_library_exports_every_declared_symbol(native):
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}
    missing = [s for s in native.declared_symbols() if s not in exported]
    assert not missing, missing
    L = native.load()
    for s in native.declared_symbols():
        assert hasattr(L, s)


def test_host_fields_bit_identical_to_oracle(native, oracle_lib):
    for pol, phi in [((0, 1, 0), (0, 0, 0)), ((0.96, 0.25, 0.0), (0.2, 1.1, 0.0))]:
        prm = make_param(HALDANE, nr=8, pol=pol, phi=phi)
        with native.Solve(prm, device=False) as s:
            t, E, A = s.fields()
        t2, E2, A2 = oracle_lib.fields(prm)
        assert np.array_equal(t, t2) and np.array_equal(E, E2) and np.array_equal(A, A2)


@pytest.mark.parametrize("path", [HMCASE0, HALDANE])
def test_host_initial_state_matches_oracle(native, oracle_lib, path):
    prm = make_param(path, nr=(10, 8))
    with native.Solve(prm, device=False) as s:
        rho = s.rho_k()
    assert np.allclose(rho, oracle_lib.rho0(path, prm, "perk"), atol=1e-13)


def test_host_initial_state_synthetic_8band(native, oracle_lib, tmp_path):
    from hhgamd import tbgen
    path = str(tmp_path / "syn8_tb.dat")
    tbgen.write_tb(path, *tbgen.synthetic(8, seed=3))
    prm = make_param(path, nr=(6, 6))
    with native.Solve(prm, device=False) as s:
        rho = s.rho_k()
    ref = oracle_lib.rho0(path, prm, "perk")
    assert np.allclose(rho, ref, atol=1e-12)
    assert np.allclose(np.einsum("kij,kjl->kil", rho, rho), rho, atol=1e-12)


def test_settings_defaults_and_nt(native):
    L = native.bind(native.load())
    st = L.Settings_new()
    sim = L.SWESim_new()
    L.SWESim_set_settings(sim, st)
    assert L.SWESim_get_nt(sim) == int(90.0 / (10.98e-3 / 0.5))  # include/settings.h:14-16
    L.SWESim_delete(sim)
    L.Settings_delete(st)


@pytest.mark.parametrize("nr,msg", [((7, 8), "even"), ((8, 5), "even"), ((2, 8), "even")])
def test_rejects_bad_grids(native, nr, msg):
    with pytest.raises(RuntimeError, match=msg):
        native.Solve(make_param(HMCASE0, nr=nr, tmax=1.0), device=False)


@pytest.mark.parametrize("tmax", [0.0, 0.02197 * 1.5])
def test_rejects_fewer_than_two_steps(native, tmax):
    """dt is t[1] - t[0] (src/solver.cpp:86-181): a solve needs nt >= 2."""
    with pytest.raises(RuntimeError, match="at least 2 time steps"):
        native.Solve(make_param(HMCASE0, nr=8, tmax=tmax, dt=0.02197), device=False)


def test_rejects_nr3(native):
    prm = make_param(HMCASE0, nr=8, tmax=1.0)
    prm["nr3"] = 2
    with pytest.raises(RuntimeError, match="nr3"):
        native.Solve(prm, device=False)


def test_missing_file_is_an_error_not_a_crash(native):
    with pytest.raises(RuntimeError, match="cannot open"):
        native.Solve(make_param("/nonexistent/x_tb.dat", nr=8, tmax=1.0), device=False)
    L = native.bind(native.load())
    assert b"cannot open" in L.hhg_last_error()


def test_garbled_file_is_an_error(native, tmp_path):
    p = tmp_path / "bad_tb.dat"
    p.write_text(open(HMCASE0).read()[:700])
    with pytest.raises(RuntimeError, match="tb file"):
        native.Solve(make_param(str(p), nr=8, tmax=1.0), device=False)


def test_run_without_device_init_is_an_error(native):
    with native.Solve(make_param(HMCASE0, nr=8, tmax=1.0), device=False) as s:
        with pytest.raises(RuntimeError, match="SWESim_init"):
            s.run()


def test_launch_queries_without_device_init_are_errors(native):
    """The launch queries describe the device run: before SWESim_init they
    fail with a message instead of guessing."""
    with native.Solve(make_param(HMCASE0, nr=8, tmax=1.0), device=False) as s:
        with pytest.raises(RuntimeError, match="SWESim_init"):
            s.kpoints_per_wave()
        with pytest.raises(RuntimeError, match="SWESim_init"):
            s.launch()


def test_bad_precision_rejected(native):
    with pytest.raises(RuntimeError, match="precision"):
        native.Solve(dict(make_param(HMCASE0, nr=8, tmax=1.0), precision="fp16"), device=False)


def test_pyswe_mirror_settings_only(native):
    from pySWE.pyswe import Settings, SWE
    prm = make_param(HMCASE0, nr=200)
    st = Settings(prm)
    assert st.nt == 4096 and st.path_lib.endswith("libwannier.so")
    bad = dict(prm, path_tb="/nonexistent_tb.dat")
    with pytest.raises(RuntimeError, match="cannot open"):
        SWE(Settings(bad))
    st.delete()


def test_server_failure_paths_without_gpu(native, tmp_path):
    """Preparation failures (missing file, footprint over the HBM budget, bad
    grid) are reported per request and never reach the device."""
    from hhgamd import serve
    with serve.Server(hbm_budget_gb=0.001, max_inflight=2, host_threads=2) as srv:
        a = srv.submit(make_param(str(tmp_path / "nope_tb.dat"), nr=8, tmax=1.0))
        b = srv.submit(make_param(HALDANE, nr=200))
        c = srv.submit(make_param(HALDANE, nr=(7, 8), tmax=1.0))
        assert srv.wait(a) == "failed" and "cannot open" in srv.error(a)
        assert srv.wait(b) == "failed" and "budget" in srv.error(b)
        assert srv.wait(c) == "failed" and "even" in srv.error(c)
        st = srv.stats()
        assert st["failed"] == 3 and st["completed"] == 0 and st["peak_inflight"] == 0
    with pytest.raises(RuntimeError, match="precision"):
        with serve.Server() as srv:
            srv.submit(dict(make_param(HALDANE, nr=8, tmax=1.0), precision="fp8"))


def test_fft2_rejects_unsupported_sizes_without_gpu(native):
    with pytest.raises(RuntimeError, match="factor"):
        native.fft2(np.zeros((17, 8), dtype=complex))
    with pytest.raises(RuntimeError, match="at most 512"):
        native.fft2(np.zeros((8, 1024), dtype=complex))


def test_fft2_plan_rejects_unsupported_sizes_without_gpu(native):
    with pytest.raises(RuntimeError, match="factor"):
        native.FFT2Plan(2, 17, 8)
    with pytest.raises(RuntimeError, match="at most 512"):
        native.FFT2Plan(1, 8, 1024)
    with pytest.raises(RuntimeError, match="nch"):
        native.FFT2Plan(0, 8, 8)


def test_library_has_no_undefined_project_symbols(native):
    out = subprocess.run(["nm", "-D", "-C", "--undefined-only", native.LIB_PATH], capture_output=True, text=True,
                         check=True)
    assert not [l for l in out.stdout.splitlines() if "hhg::" in l]


def test_standalone_driver_reports_errors(native, tmp_path):
    exe = os.path.join(PKG, "hhg_swe")
    assert os.path.exists(exe)
    r = subprocess.run([exe, str(tmp_path / "missing_tb.dat"), "nr1=8", "nr2=8"], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr
    r = subprocess.run([exe, HMCASE0, "bogus=1"], capture_output=True, text=True)
    assert r.returncode == 2 and "unknown argument" in r.stderr


def test_test_files_needs_device_init(native):
    with native.Solve(make_param(HMCASE0, nr=8, tmax=1.0), device=False) as s:
        s.lib.SWESim_test_files(s.ptr)
        assert s.lib.SWESim_get_status(s.ptr) != 0
        assert b"SWESim_init" in s.lib.SWESim_get_error(s.ptr)


def test_server_rejects_bad_policy(native):
    from hhgamd import serve
    with pytest.raises(RuntimeError, match="admission policy"):
        serve.Server(policy="random")
    with pytest.raises(RuntimeError, match="aging"):
        serve.Server(aging_s=0.0)


def test_autotune_switch_is_shared_across_threads(native):
    """The process-wide tuning switch set in one host thread is seen by the others."""
    import threading
    seen = []
    prev = native.get_autotune()
    native.set_autotune(True)
    try:
        t = threading.Thread(target=lambda: seen.append(native.get_autotune()))
        t.start()
        t.join(timeout=30)
        assert seen == [True]
    finally:
        native.set_autotune(prev)
    assert native.get_autotune() == prev
