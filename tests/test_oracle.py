"""Pins of the CPU oracle (oracle/swe_oracle.cpp) -- no GPU.

Each check is chosen so that a plausible mistake in the restatement (a
transposed block, a wrong FFT sign or shift, a dropped unit factor, a wrong
RK4 weight, a wrong window) fails it. Sources of truth: closed forms,
numpy.fft / numpy.linalg, libm through ctypes, an independent numpy
restatement (oracle/fftform_np.py), and the committed golden fixtures.
"""
import ctypes as ct

import numpy as np
import pytest

from conftest import HALDANE, HMCASE0, GOLDEN, make_param, rel_l2

HA = 27.2114          # eV per Hartree (include/constants.h:10-11)
BOHR = 0.5291772      # Angstrom per bohr (include/constants.h:6-7)


# ---------------------------------------------------------------- parser ----
def test_parser_hmcase0_by_inspection(oracle_lib):
    """hmcase0_tb.dat lines 9-91: on-site -2.1/+2.1, NN -2.42, B at x = 1.42028 A."""
    lat, R, H, r = oracle_lib.model(HMCASE0)
    assert np.allclose(lat[0], [2.130422493, -1.23, 0]) and np.allclose(lat[2], [0, 0, 15])
    i0 = [tuple(x) for x in R].index((0, 0, 0))
    assert np.allclose(H[i0], [[-2.1, -2.42], [-2.42, 2.1]])
    # file line "1 2 -2.42" at R = (-1,0,0) is stored at [1][0] (src/wannier_tb.cpp:73-75)
    im = [tuple(x) for x in R].index((-1, 0, 0))
    assert H[im][1, 0] == -2.42 and H[im][0, 1] == 0
    ip = [tuple(x) for x in R].index((1, 0, 0))
    assert H[ip][0, 1] == -2.42 and H[ip][1, 0] == 0
    assert np.isclose(r[0, i0, 1, 1].real, 1.42028) and abs(r[0, i0, 0, 0]) < 1e-12
    assert np.all(r[1] == 0) and np.all(r[2] == 0)


def test_parser_matches_numpy_parser(oracle_lib):
    from oracle import fftform_np
    for path in (HMCASE0, HALDANE):
        lat, R, H, r = oracle_lib.model(path)
        lat2, R2, H2, r2 = fftform_np.parse_tb(path)
        assert np.array_equal(lat, lat2) and np.array_equal(R, R2)
        assert np.array_equal(H, H2) and np.array_equal(r, r2)


# ------------------------------------------------------------------- FFT ----
@pytest.mark.parametrize("shape", [(4, 4), (6, 10), (12, 20), (30, 30), (8, 50), (14, 22)])
def test_fft2_vs_numpy(oracle_lib, shape):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert np.allclose(oracle_lib.fft2(a, -1), np.fft.fft2(a), atol=1e-12)
    assert np.allclose(oracle_lib.fft2(a, +1), np.fft.ifft2(a) * a.size, atol=1e-12)


def test_fft2_brute_force_dft(oracle_lib):
    rng = np.random.default_rng(2)
    a = rng.standard_normal((6, 4)) + 1j * rng.standard_normal((6, 4))
    j1, j2 = np.meshgrid(np.arange(6), np.arange(4), indexing="ij")
    ref = np.array([[np.sum(a * np.exp(-2j * np.pi * (m1 * j1 / 6 + m2 * j2 / 4))) for m2 in range(4)]
                    for m1 in range(6)])
    assert np.allclose(oracle_lib.fft2(a), ref, atol=1e-12)


# ------------------------------------------------------------ eigensolver ----
@pytest.mark.parametrize("n", [2, 3, 8, 16])
def test_jacobi_eigh_vs_numpy(oracle_lib, n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = A + A.conj().T
    w, U = oracle_lib.eigh(A)
    assert np.allclose(np.sort(w), np.linalg.eigvalsh(A), atol=1e-12)
    assert np.allclose(U.conj().T @ A @ U, np.diag(w), atol=1e-11)
    assert np.allclose(U.conj().T @ U, np.eye(n), atol=1e-12)


def test_equilibrium_rho_is_negative_band_projector(oracle_lib):
    rng = np.random.default_rng(5)
    A = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    A = A + A.conj().T
    w, V = np.linalg.eigh(A)
    P = V[:, w < 0] @ V[:, w < 0].conj().T
    assert np.allclose(oracle_lib.equilibrium_rho(A), P, atol=1e-12)


# ------------------------------------------------------------- H(k) bands ----
def test_hmcase0_bands_closed_form(oracle_lib):
    """Honeycomb: E(k) = +-sqrt(2.1^2 + 2.42^2 |1 + e^{-i k.a1} + e^{-i k.a2}|^2) eV."""
    N1, N2 = 10, 12
    hk = oracle_lib.hk(HMCASE0, make_param(HMCASE0, nr=(N1, N2)))
    m1, m2 = np.meshgrid(np.arange(N1), np.arange(N2), indexing="ij")
    f = 1 + np.exp(-2j * np.pi * m1 / N1) + np.exp(-2j * np.pi * m2 / N2)
    e = np.sqrt(2.1 ** 2 + 2.42 ** 2 * np.abs(f) ** 2).reshape(-1) / HA
    w = np.linalg.eigvalsh(hk)
    assert np.allclose(w[:, 0], -e, atol=1e-14) and np.allclose(w[:, 1], e, atol=1e-14)
    # off-diagonal element sign/orientation: H_k[0][1] = -2.42 (1 + e^{-ik.a1} + e^{-ik.a2}) / Ha
    assert np.allclose(hk[:, 0, 1], (-2.42 * f / HA).reshape(-1), atol=1e-15)


def test_haldane_gap_closed_form(oracle_lib):
    """Haldane gap 2|M - 3 sqrt(3) t2 sin(phi)| at the Dirac points (30 | 3, so K is on the grid)."""
    hk = oracle_lib.hk(HALDANE, make_param(HALDANE, nr=30))
    w = np.linalg.eigvalsh(hk) * HA
    assert np.isclose((w[:, 1] - w[:, 0]).min(), 2 * abs(0.3 - 3 * np.sqrt(3) * 0.1), atol=1e-12)


def test_initial_state_projector_invariants(oracle_lib):
    for path in (HMCASE0, HALDANE):
        prm = make_param(path, nr=(8, 6))
        hk = oracle_lib.hk(path, prm)
        for mode in ("perk", "fft"):
            rho = oracle_lib.rho0(path, prm, mode)
            assert np.allclose(rho @ rho, rho, atol=1e-13)
            assert np.allclose(np.trace(rho, axis1=1, axis2=2), 1.0, atol=1e-13)
            assert np.allclose(rho, rho.conj().transpose(0, 2, 1), atol=1e-14)
            assert np.allclose(hk @ rho - rho @ hk, 0, atol=1e-14)
        assert np.allclose(oracle_lib.rho0(path, prm, "perk"), oracle_lib.rho0(path, prm, "fft"), atol=1e-13)


# ----------------------------------------------------------------- fields ----
def test_fields_readme_closed_forms(oracle_lib):
    """README pulse: nt = int(90/0.02197) = 4096, dt = 0.02197 fs in au, T = lambda/c."""
    prm = make_param(HMCASE0, nr=10)
    t, E, A = oracle_lib.fields(prm)
    assert len(t) == 4096
    assert np.isclose(t[1], 0.02197 / 0.02418884, rtol=1e-15)
    libm = ct.CDLL("libm.so.6")
    libm.sqrtf.argtypes = [ct.c_float]
    libm.sqrtf.restype = ct.c_float
    E0 = float(libm.sqrtf(1e12 / 3.56e16))
    assert np.isclose(E0, 5.29999e-3, rtol=1e-5)
    assert np.abs(E[1]).max() <= E0 and np.abs(E[1]).max() > 0.99 * E0
    assert np.all(E[0] == 0) and np.all(E[2] == 0)
    tf = 80.0 / 0.02418884
    assert np.all(E[1][t >= tf] == 0.0)
    # A is the left Riemann sum of -E (src/efield.cpp:45-49)
    dt = t[1] - t[0]
    assert np.allclose(A[1], -np.concatenate([[0.0], np.cumsum(E[1][:-1]) * dt]), atol=1e-15)
    # carrier period 2 pi / omega = lambda / c = 30000 A / 2997.025 A/fs
    period = 30000.0 / 2997.025 / 0.02418884
    zeros = t[1:][(np.sign(E[1][1:]) != np.sign(E[1][:-1])) & (t[1:] < tf - 10)]
    assert np.isclose(np.median(np.diff(zeros)) * 2, period, rtol=2e-3)


def test_fields_bit_identical_to_libm_restatement(oracle_lib):
    from oracle import fftform_np
    for pol, phi in [((0, 1, 0), (0, 0, 0)), ((0.6, 0.8, 0.1), (0.3, -1.2, 2.0))]:
        prm = make_param(HMCASE0, nr=10, tmax=20.0, pol=pol, phi=phi, intensity=3e12, tmax_field=15.0)
        t, E, A = oracle_lib.fields(prm)
        t2, E2, A2 = fftform_np.fields(prm)
        assert np.array_equal(t, t2) and np.array_equal(E, E2) and np.array_equal(A, A2)


def test_fields_golden(oracle_lib):
    g = np.load(f"{GOLDEN}/fields_readme.npz")
    t, E, A = oracle_lib.fields(make_param(HMCASE0, nr=10))
    assert np.array_equal(t, g["t"]) and np.array_equal(E, g["E"]) and np.array_equal(A, g["A"])


# --------------------------------------------------------------- velocity ----
def test_velocity_direct_vs_fft(oracle_lib):
    for path in (HMCASE0, HALDANE):
        prm = make_param(path, nr=(8, 10))
        vd = oracle_lib.velocity_window(path, prm, "direct")
        vf = oracle_lib.velocity_window(path, prm, "fft")
        assert np.allclose(vd, vf, atol=1e-13)


def test_velocity_closed_form_hmcase0(oracle_lib):
    """Positions only at R = 0 and diagonal: v_x,ij(R) = i (R_x + x_i - x_j) H_ij(R), in eV*A."""
    lat, R, H, r = oracle_lib.model(HMCASE0)
    v = oracle_lib.velocity_window(HMCASE0, make_param(HMCASE0, nr=10), "direct")
    x = np.real(np.diag(r[0, [tuple(q) for q in R].index((0, 0, 0))]))
    for w1 in range(4):
        for w2 in range(4):
            n1, n2 = w1 - 2, w2 - 2
            key = (n1, n2, 0)
            Hn = H[[tuple(q) for q in R].index(key)] if key in [tuple(q) for q in R] else np.zeros((2, 2))
            for a in range(2):
                Ra = n1 * lat[0, a] + n2 * lat[1, a]
                ref = 1j * (Ra + (x[:, None] - x[None, :]) * (a == 0)) * Hn
                assert np.allclose(v[a, w1, w2], ref, atol=1e-12)
    assert np.all(v[2] == 0)


# ---------------------------------------------------------- time loop, J ----
def test_fftform_cpp_vs_numpy_restatement(oracle_lib):
    from oracle import fftform_np
    for path, pol, phi in [(HMCASE0, (0.6, 0.8, 0), (0.1, 0.3, 0)), (HALDANE, (1, 0, 0), (0, 0, 0))]:
        prm = make_param(path, nr=(6, 8), tmax=4.0, tmax_field=3.0, intensity=2e13, pol=pol, phi=phi)
        _, Jn = fftform_np.run(path, prm)
        _, Jc = oracle_lib.run(path, prm, mode="fft")
        assert rel_l2(Jc, Jn) < 1e-12


@pytest.mark.parametrize("path", [HMCASE0, HALDANE])
def test_perk_vs_fftform_full_pulse(oracle_lib, path):
    prm = make_param(path, nr=10, pol=(0.6, 0.8, 0.0), phi=(0.0, 0.4, 0.0))
    _, Jf = oracle_lib.run(path, prm, mode="fft")
    _, Jk = oracle_lib.run(path, prm, mode="perk", nthreads=4)
    assert rel_l2(Jk, Jf) < 1e-11


def test_perk_thread_count_invariance(oracle_lib):
    prm = make_param(HALDANE, nr=(8, 12), tmax=20.0)
    _, J1 = oracle_lib.run(HALDANE, prm, mode="perk", nthreads=1)
    _, J5 = oracle_lib.run(HALDANE, prm, mode="perk", nthreads=5)
    assert rel_l2(J5, J1) < 1e-12  # only the k-partition of the sum differs


def window_current(path, prm, rho, ti):
    """J_a[ti] = (1/N) sum_k sum_{n in 4x4 window} sum_ij v_a(n)_ij conj(rho_k,ij) e^{-i(k - A[ti]).R(n)}
    from final states rho (N, nb, nb): the k-space form of src/observable.cpp:16-30
    (v in eV*Angstrom, R in bohr, k.R(n) = 2 pi (m1 n1 / N1 + m2 n2 / N2))."""
    import oracle
    lat, _, _, _ = oracle.model(path)
    _, _, A = oracle.fields(prm)
    v = oracle.velocity_window(path, prm)  # (3, 4, 4, nb, nb), n1, n2 in {-2..1}
    N1, N2 = prm["nr1"], prm["nr2"]
    m1, m2 = np.divmod(np.arange(N1 * N2), N2)
    J = np.zeros(3, dtype=complex)
    for i1, n1 in enumerate(range(-2, 2)):
        for i2, n2 in enumerate(range(-2, 2)):
            R = (n1 * lat[0] + n2 * lat[1]) / BOHR
            ph = np.exp(-2j * np.pi * (m1 * n1 / N1 + m2 * n2 / N2)) * np.exp(1j * A[:, ti] @ R)
            for a in range(3):
                J[a] += np.sum(ph * np.einsum("ij,kij->k", v[a, i1, i2], np.conj(rho)))
    return J / (N1 * N2)


def test_window_current_is_the_closed_form_velocity_for_hmcase0(oracle_lib):
    """hmcase0's hoppings all lie inside the 4x4 window, so the windowed current
    of src/observable.cpp:16-30 must equal (1/N) sum_k Tr(v(k - A) rho_k) with
    the physical velocity in closed form (conftest.hmcase0_closed_form_current);
    a wrong window, phase sign, unit or orbital-offset term breaks this."""
    from conftest import hmcase0_closed_form_current
    prm = make_param(HMCASE0, nr=(8, 10), tmax=0.02197 * 70 + 1e-9, tmax_field=1.4, intensity=5e12,
                     pol=(0.6, 0.8, 0.0), phi=(0.0, 0.3, 0.0))
    nt = oracle_lib.nt(prm)
    rho = oracle_lib.propagate_k(HMCASE0, prm, np.arange(80), nthreads=4)
    _, _, A = oracle_lib.fields(prm)
    _, J = oracle_lib.run(HMCASE0, prm, mode="perk", nthreads=4)
    Jc = hmcase0_closed_form_current(rho, A[:, nt - 1], 8, 10)
    assert np.abs(J[nt - 1, 0]) > 1e-6 and np.abs(J[nt - 1, 1]) > 1e-6
    assert np.abs(Jc - J[nt - 1]).max() <= 1e-12 * np.abs(J[nt - 1]).max()
    # the closed form sees the field: with A = 0 it differs
    assert np.abs(hmcase0_closed_form_current(rho, np.zeros(3), 8, 10) - J[nt - 1]).max() > 1e-6 * np.abs(J).max()


def test_propagate_k_closes_the_last_current(oracle_lib, tmp_path):
    """The final states of oracle_propagate_k over every k reproduce the last
    current sample of oracle_run through the window's k-space closed form,
    and stay Hermitian with their initial traces."""
    from hhgamd import tbgen
    path = str(tmp_path / "syn4_tb.dat")
    tbgen.write_tb(path, *tbgen.synthetic(4, seed=7))
    prm = make_param(path, nr=(6, 8), tmax=0.02197 * 24 + 1e-9, tmax_field=0.5, intensity=5e11,
                     pol=(0.6, 0.8, 0.0))
    nt = oracle_lib.nt(prm)
    rho = oracle_lib.propagate_k(path, prm, np.arange(48), nthreads=3)
    rho0 = oracle_lib.rho0(path, prm)
    assert np.abs(rho - np.conj(np.transpose(rho, (0, 2, 1)))).max() < 1e-14
    assert np.abs(np.trace(rho, axis1=1, axis2=2) - np.trace(rho0, axis1=1, axis2=2)).max() < 1e-13
    assert np.abs(rho - rho0).max() > 1e-6  # the field moved the states
    _, J = oracle_lib.run(path, prm, mode="perk", nthreads=2)
    Jw = window_current(path, prm, rho, nt - 1)
    assert np.abs(Jw - J[nt - 1]).max() <= 1e-12 * np.abs(J[nt - 1]).max()
    # a subset in another order gives the same states
    sub = np.array([47, 3, 20], dtype=np.int32)
    assert np.array_equal(oracle_lib.propagate_k(path, prm, sub, nthreads=2), rho[sub])


@pytest.mark.parametrize("name,path,n", [("hmcase0", HMCASE0, 10), ("hmcase0", HMCASE0, 30), ("haldane", HALDANE, 30)])
def test_oracle_reproduces_golden(oracle_lib, name, path, n):
    g = np.load(f"{GOLDEN}/golden_{name}_{n}.npz")
    t, J = oracle_lib.run(path, make_param(path, nr=n), mode="perk")
    assert np.array_equal(t, g["t"])
    assert rel_l2(J, g["J"]) < 1e-11


def test_field_free_state_is_stationary(oracle_lib):
    """E = 0, A = 0: the T=0 projector commutes with H(k), so J(t) stays J(0)."""
    prm = make_param(HALDANE, nr=8, tmax=10.0, intensity=0.0)
    _, J = oracle_lib.run(HALDANE, prm, mode="perk", nthreads=2)
    assert np.abs(J - J[0]).max() < 1e-15


def test_rk4_amplification_after_pulse(oracle_lib):
    """After the pulse, H(k - A_end) is constant and every coherence of the
    2-band system advances per step by the RK4 factor g(z) = 1 + z + z^2/2 +
    z^3/6 + z^4/24 with z = -i w dt, w = E+ - E- of the honeycomb closed form
    at k - A_end. J(t) must lie in the span of {g(-i w_k dt)^n,
    g(+i w_k dt)^n, 1}; a wrong stage weight or combination breaks that."""
    N1 = N2 = 4
    prm = make_param(HMCASE0, nr=N1, tmax=6.0, tmax_field=3.0, intensity=5e13, pol=(0.6, 0.8, 0.0))
    t, J = oracle_lib.run(HMCASE0, prm, mode="perk", nthreads=1)
    _, E, A = oracle_lib.fields(prm)
    dt = t[1] - t[0]
    tf = 3.0 / 0.02418884
    i0 = int(np.searchsorted(t, tf)) + 2
    Aend = A[:, i0]
    assert np.all(E[:, i0 - 1:] == 0)
    lat, _, _, _ = oracle_lib.model(HMCASE0)
    a1, a2 = lat[0] / BOHR, lat[1] / BOHR
    m1, m2 = np.meshgrid(np.arange(N1), np.arange(N2), indexing="ij")
    # q.a_d = 2 pi m_d / N_d - A.a_d
    f = 1 + np.exp(-1j * (2 * np.pi * m1 / N1 - Aend @ a1)) + np.exp(-1j * (2 * np.pi * m2 / N2 - Aend @ a2))
    w = 2 * np.sqrt(2.1 ** 2 + 2.42 ** 2 * np.abs(f.reshape(-1)) ** 2) / HA
    g = lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    n = np.arange(len(t) - i0)
    basis = [np.ones_like(n, dtype=complex)]
    for wk in np.unique(np.round(w, 14)):
        basis += [g(-1j * wk * dt) ** n, g(1j * wk * dt) ** n]
    B = np.stack(basis, axis=1)
    for a in range(2):
        y = J[i0:, a]
        coef, *_ = np.linalg.lstsq(B, y, rcond=None)
        assert np.linalg.norm(B @ coef - y) / np.linalg.norm(y) < 1e-12
    # the same fit with the textbook weights perturbed fails
    gbad = lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 20
    Bb = np.stack([np.ones_like(n, dtype=complex)] + sum(
        [[gbad(-1j * wk * dt) ** n, gbad(1j * wk * dt) ** n] for wk in np.unique(np.round(w, 14))], []), axis=1)
    y = J[i0:, 1]
    coef, *_ = np.linalg.lstsq(Bb, y, rcond=None)
    assert np.linalg.norm(Bb @ coef - y) / np.linalg.norm(y) > 1e-8  # measured: 1.5e-7 (correct: 1.5e-15)


def test_oracle_rejects_bad_grids(oracle_lib):
    for nr in [(7, 8), (8, 9), (2, 8)]:
        with pytest.raises(RuntimeError, match="even"):
            oracle_lib.run(HMCASE0, make_param(HMCASE0, nr=nr, tmax=1.0))
    with pytest.raises(RuntimeError, match="cannot open"):
        oracle_lib.run("/nonexistent_tb.dat", make_param(HMCASE0, nr=8, tmax=1.0))
