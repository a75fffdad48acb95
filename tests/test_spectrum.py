"""Spectra and per-harmonic intensities (hhgamd/spectrum.py) on signals whose
spectra are known in closed form (README.md:202-210, scripts/pol_scan.py:31-57):
a sum of harmonics that fit the record an integer number of times puts each
harmonic into one FFT bin with |rfft|^2 = (amplitude * nt / 2)^2."""
import numpy as np

from hhgamd import spectrum

LAM = 3000.0


def test_carrier_frequency_closed_form():
    # omega = 2 pi c / lambda, c = 2997.025 Angstrom/fs, 3000 nm, in 1/au (0.02418884 fs)
    assert np.isclose(spectrum.carrier_omega_au(LAM), 2 * np.pi * 2997.025 / 30000.0 * 0.02418884, rtol=1e-15)
    assert np.isclose(spectrum.carrier_omega_au(LAM), 0.015183, atol=5e-7)


def _signal(amps, periods=8, per_period=256):
    """J(t) = sum_q a_q sin(q w0 t) over `periods` carrier periods, so harmonic q sits in bin q * periods."""
    w0 = spectrum.carrier_omega_au(LAM)
    nt = periods * per_period
    dt = 2 * np.pi / w0 / per_period
    t = np.arange(nt) * dt
    J = sum(a * np.sin(q * w0 * t) for q, a in amps.items())
    return t, J, w0, dt, nt


def test_readme_spectrum_bins():
    amps = {1: 1.0, 3: 0.25, 5: 1e-3, 9: 2e-5}
    t, J, w0, dt, nt = _signal(amps)
    f, P = spectrum.readme_spectrum(t, J + 1j * 7.0)  # the imaginary part is ignored
    for q, a in amps.items():
        b = q * 8
        assert np.isclose(f[b], q * w0 / (2 * np.pi), rtol=1e-12)
        assert np.isclose(P[b], (a * nt / 2) ** 2, rtol=1e-9)
    other = np.ones(len(P), dtype=bool)
    other[[q * 8 for q in amps]] = False
    assert P[other].max() < 1e-18 * P.max()


def test_harmonic_intensities_closed_form():
    """dJ/dt by np.gradient: central differences scale harmonic q by sin(q w0 dt)/(q w0 dt)
    (the end points use one-sided differences: O(1/nt) leakage)."""
    amps = {1: 1.0, 3: 0.3, 5: 0.02, 7: 4e-3, 11: 1e-4}
    t, J, w0, dt, nt = _signal(amps, periods=16, per_period=512)
    q, I = spectrum.harmonic_intensities(t, J, LAM, qmax=15)
    assert np.array_equal(q, np.arange(1, 16))
    for k, a in amps.items():
        ref = (a * np.sin(k * w0 * dt) / dt * nt / 2) ** 2
        assert abs(I[k - 1] / ref - 1) < 2e-3, k
    for k in set(range(1, 16)) - set(amps):  # only end-point leakage between the harmonics
        assert I[k - 1] < 1e-6 * I[0]
    assert spectrum.cutoff_order(q, I) == 11          # 1e-4^2 * 11^2 / 1 ~ 1.2e-6 of the first
    assert spectrum.cutoff_order(q, I, dynamic_range=1e-4) == 7


def test_polscan_spectrum_of_a_harmonic():
    """fftshift(fft(ifftshift(gradient(J)))) of a carrier cosine peaks at freq/w0 = +-1 (c = 300 nm/fs)."""
    lam, dt, periods, ppp = 1600.0, 0.01, 12, 200
    w0 = 2 * np.pi * 300 / lam   # 1/fs, scripts/pol_scan.py:57
    tmax = periods * 2 * np.pi / w0
    nt = int(tmax / dt)
    t = np.arange(nt) * dt
    J = np.cos(w0 * t) * np.sin(np.pi * t / tmax) ** 2
    qax, hhg = spectrum.polscan_spectrum(J, dt, tmax, lam)
    assert len(qax) == nt and np.isclose(qax[0], -qax[-1])
    peak = np.argsort(np.abs(hhg))[-2:]
    assert np.allclose(np.sort(np.abs(qax[peak])), [1.0, 1.0], atol=2 * (qax[1] - qax[0]))
