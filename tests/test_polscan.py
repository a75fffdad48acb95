"""Polarisation-scan driver (scripts/pol_scan.py): angle grid, per-angle
parameters and spectra on the CPU; on the GPU the batched and served scans
equal the angles run one by one."""
import numpy as np
import pytest

from conftest import HALDANE, make_param


def test_angles_and_parameters_follow_the_reference():
    from hhgamd import polscan
    a = polscan.angles(15)
    assert len(a) == 15 and a[0] == 0.0 and np.isclose(a[-1], 2 * np.pi)  # scripts/pol_scan.py:46
    ps = polscan.params_for(make_param(HALDANE, nr=8), a[:3])
    for p, x in zip(ps, a[:3]):
        assert np.array_equal(p["pol_vec"], np.array([np.sin(x), np.cos(x), 0.0]))  # :24


def test_spectra_restate_pol_scan():
    """fftshift(fft(ifftshift(gradient(J, dt)))) per axis and freq / w0 with
    c = 300, restated here from scripts/pol_scan.py:31-57 on a test signal."""
    from hhgamd import polscan
    prm = make_param(HALDANE, nr=8, tmax=9.0, dt=0.05, lam=1600.0)
    nt = int(prm["tmax"] / prm["dt"])
    t = np.arange(nt) * prm["dt"]
    rng = np.random.default_rng(1)
    J = np.stack([np.sin((a + 1) * t) * np.exp(-t / 5) + 1j * rng.normal(size=nt) * 1e-3 for a in range(3)], axis=1)
    freq, w0, hhg = polscan.spectra(J, prm)
    w0_ref = 2.0 * np.pi * 300 / 1600.0
    dw = 2.0 * np.pi / prm["tmax"]
    freq_ref = np.linspace(-nt * dw / 2.0, nt * dw / 2.0, nt) / w0_ref
    assert w0 == w0_ref and np.array_equal(freq, freq_ref)
    for a in range(3):
        acc = np.gradient(J[:, a], prm["dt"])
        assert np.array_equal(hhg[a], np.fft.fftshift(np.fft.fft(np.fft.ifftshift(acc))))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["batch", "serve"])
def test_scan_equals_angles_alone(mode):
    from hhgamd import lib, polscan
    prm = make_param(HALDANE, nr=(12, 16), tmax=8.0, tmax_field=7.0, intensity=2e12)
    rec = polscan.run(prm, n_angles=4, mode=mode)
    assert rec["data_sim"].shape == (4, 3, int(prm["tmax"] / prm["dt"]))
    for i, p in enumerate(polscan.params_for(prm, rec["pol_angles"])):
        with lib.Solve(p) as s:
            s.run()
            J = s.current()[1]
        _, _, hhg = polscan.spectra(J, p)
        assert np.array_equal(rec["data_sim"][i], hhg)  # every angle keeps its own spectra
    assert not np.allclose(rec["data_sim"][0], rec["data_sim"][1])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["batch", "serve"])
def test_scan_spectra_vs_oracle(oracle_lib, mode):
    """Every angle's spectra of a batched or served scan against the spectra of
    the oracle's J(t) for the same angle (scripts/pol_scan.py:31-57 applied to
    the FP64 oracle); the scan's currents are FP64, so the spectra agree to the
    J parity bar."""
    from conftest import HMCASE0
    from hhgamd import polscan
    prm = make_param(HMCASE0, nr=(12, 14), tmax=12.0, tmax_field=10.0, intensity=3e12)
    rec = polscan.run(prm, n_angles=5, mode=mode)
    for i, p in enumerate(polscan.params_for(prm, rec["pol_angles"])):
        _, Jref = oracle_lib.run(HMCASE0, p, mode="perk")
        _, _, ref = polscan.spectra(Jref, p)
        scale = max(np.linalg.norm(ref[0]), np.linalg.norm(ref[1]))  # an axis may vanish by symmetry
        for a in range(3):
            err = np.linalg.norm(rec["data_sim"][i, a] - ref[a]) / scale
            assert err <= 1e-10, (i, a, err)
