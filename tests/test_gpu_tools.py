"""Standalone driver (src/main.cpp counterpart) and the SWESim_test_files
dumps (src/swe.cpp:131-181): contents against numpy restatements."""
import os
import subprocess

import numpy as np
import pytest

from conftest import HMCASE0, PKG, make_param

pytestmark = pytest.mark.gpu


def test_driver_current_equals_library(tmp_path):
    from hhgamd import lib
    out = tmp_path / "j.dat"
    r = subprocess.run([os.path.join(PKG, "hhg_swe"), HMCASE0, "nr1=10", "nr2=12", "tmax=10", "dt=0.02197", "tmax_field=8",
                        "intensity=2e12", "pol=0.6,0.8,0", f"out={out}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    d = np.loadtxt(out)
    p = make_param(HMCASE0, nr=(10, 12), tmax=10.0, tmax_field=8.0, intensity=2e12, pol=(0.6, 0.8, 0.0))
    with lib.Solve(p) as s:
        s.run()
        t, J = s.current()
    assert np.array_equal(d[:, 0], t)
    assert np.array_equal(d[:, 1::2] + 1j * d[:, 2::2], J)


def _matrix(path, n, nb2):
    with open(path) as f:
        tag = f.readline().strip()
    a = np.loadtxt(path, skiprows=1)
    assert a.shape == (n * nb2, 3)
    return tag, (a[:, 1] * np.exp(1j * a[:, 2])).reshape(n, nb2)


def test_test_files_contents(tmp_path):
    from hhgamd import lib
    N1, N2 = 8, 10
    p = make_param(HMCASE0, nr=(N1, N2), tmax=2.0, tmax_field=1.5, pol=(0.6, 0.8, 0.0))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        with lib.Solve(p) as s:
            s.lib.SWESim_test_files(s.ptr)
            s._check("test_files")
            t, E, A = s.fields()
        with lib.Solve(p, device=False) as s:
            rho0 = s.rho_k().reshape(N1 * N2, -1)
    finally:
        os.chdir(cwd)
    T = tmp_path / "test"
    # grid: R = n1 a1 + n2 a2, n = i - N/2; k = (m1/N1) b1 + (m2/N2) b2
    lat = np.array([[2.130422493, -1.23, 0.0], [2.130422493, 1.23, 0.0], [0.0, 0.0, 15.0]])
    b = 2 * np.pi * np.linalg.inv(lat).T
    n1, n2 = np.meshgrid(np.arange(N1) - N1 // 2, np.arange(N2) - N2 // 2, indexing="ij")
    R = n1.reshape(-1, 1) * lat[0] + n2.reshape(-1, 1) * lat[1]
    m1, m2 = np.meshgrid(np.arange(N1), np.arange(N2), indexing="ij")
    K = (m1.reshape(-1, 1) / N1) * b[0] + (m2.reshape(-1, 1) / N2) * b[1]
    assert np.allclose(np.loadtxt(T / "Rvecs_SI.dat", skiprows=1), R, rtol=1e-5, atol=1e-5)
    assert np.allclose(np.loadtxt(T / "Kvecs_SI.dat", skiprows=1), K, rtol=1e-5, atol=1e-5)
    assert np.allclose(np.loadtxt(T / "Rvecs_AU.dat", skiprows=1), R / 0.5291772, rtol=1e-5, atol=1e-5)
    # H(R) -> H_k = FFT(ifftshift(H)) and the round trip
    tag, H = _matrix(T / "hamiltonian_AU.dat", N1 * N2, 4)
    assert tag == "_AU_"
    Hg = H.reshape(N1, N2, 4)
    Hk_ref = np.fft.fft2(np.fft.ifftshift(Hg, axes=(0, 1)), axes=(0, 1)).reshape(-1, 4)
    _, Hk = _matrix(T / "hamiltonian_k_AU.dat", N1 * N2, 4)
    assert np.allclose(Hk, Hk_ref, rtol=1e-5, atol=1e-6)
    _, Hr = _matrix(T / "hamiltonian_r_AU.dat", N1 * N2, 4)
    assert np.allclose(Hr, H, rtol=1e-5, atol=1e-6)
    # T = 0 state in the W gauge equals the solve's initial state; band basis unitary
    _, RW = _matrix(T / "rho_k_W_AU.dat", N1 * N2, 4)
    assert np.allclose(RW, rho0, atol=1e-5)
    _, U = _matrix(T / "Uk_AU.dat", N1 * N2, 4)
    Um = U.reshape(-1, 2, 2)
    assert np.allclose(np.einsum("kji,kjl->kil", Um.conj(), Um), np.eye(2), atol=1e-5)
    _, HH = _matrix(T / "hamiltonian_k_H_AU.dat", N1 * N2, 4)
    assert np.allclose(HH[:, 1], 0, atol=1e-5) and np.all(HH.reshape(-1, 2, 2)[:, 0, 0].real <= HH.reshape(-1, 2, 2)[:, 1, 1].real + 1e-6)
    # fields in au
    fe = np.loadtxt(tmp_path / "fields" / "efield.dat", skiprows=1)
    assert np.allclose(fe[:, 0], t, rtol=1e-5) and np.allclose(fe[:, 1:], E.T, rtol=1e-5, atol=1e-12)
    for name in ("x_bc.dat", "y_bc.dat", "z_bc.dat", "rho_k_H_AU.dat", "rho_r_W_AU.dat", "hamiltonian_SI.dat",
                 "Kvecs_AU.dat", "hamiltonian_k_W_AU.dat"):
        assert (T / name).exists()
