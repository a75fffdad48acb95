"""HHG spectra and per-harmonic intensities from J(t), in FP64.

Two spectrum definitions of the reference:
  * ``readme_spectrum``: |rfft(J.real)|^2 over rfftfreq(nt, dt) with t in au
    (README.md:202-210);
  * ``polscan_spectrum``: fftshift(fft(ifftshift(gradient(J, dt)))) with dt and
    the frequency axis in the API's time units and the harmonic axis freq/w0,
    w0 = 2 pi c / lambda with c = 300 nm/fs (scripts/pol_scan.py:31-57).

``harmonic_intensities`` integrates the power spectrum of the dipole
acceleration dJ/dt over windows [q - 1/2, q + 1/2] in units of the carrier
frequency (the frequency of src/efield.cpp:10-11 converted to au), which is
the per-harmonic figure compared between precisions.
"""
import numpy as np

C_SI = 2997.025          # Angstrom / fs (include/constants.h:16)
FS_PER_AU = 0.02418884   # include/constants.h:8-9


def carrier_omega_au(lambda_nm):
    """omega = 2 pi c / lambda of src/efield.cpp:10-11, converted to 1/au (src/efield.cpp:72)."""
    w = 2.0 * np.pi * C_SI / (lambda_nm * 10.0)
    return w * (1.0 / (1.0 / FS_PER_AU))


def readme_spectrum(t, j):
    dt = t[1] - t[0]
    return np.fft.rfftfreq(len(t), d=dt), np.abs(np.fft.rfft(np.real(j))) ** 2


def polscan_spectrum(j, dt, tmax, lambda_nm):
    """Returns (harmonic axis freq/w0, complex spectrum) as scripts/pol_scan.py:23-57."""
    acc = np.gradient(j, dt)
    hhg = np.fft.fftshift(np.fft.fft(np.fft.ifftshift(acc)))
    nw = int(tmax / dt)
    dw = 2.0 * np.pi / tmax
    wmax = nw * dw / 2.0
    freq = np.linspace(-wmax, wmax, nw)
    w0 = 2.0 * np.pi * 300 / lambda_nm
    return freq / w0, hhg


def acceleration_power(t, j):
    """Power spectrum |FFT(dJ/dt)|^2 on the one-sided angular-frequency axis (1/au)."""
    dt = t[1] - t[0]
    acc = np.gradient(np.real(j), dt)
    spec = np.abs(np.fft.rfft(acc)) ** 2
    omega = 2.0 * np.pi * np.fft.rfftfreq(len(t), d=dt)
    return omega, spec


def harmonic_intensities(t, j, lambda_nm, qmax=None):
    """I_q = sum of |FFT(dJ/dt)|^2 over omega / omega_0 in [q - 1/2, q + 1/2), q = 1..qmax."""
    omega, spec = acceleration_power(t, j)
    order = omega / carrier_omega_au(lambda_nm)
    if qmax is None:
        qmax = int(order.max() - 0.5)
    q = np.arange(1, qmax + 1)
    I = np.array([spec[(order >= k - 0.5) & (order < k + 0.5)].sum() for k in q])
    return q, I


def cutoff_order(q, I, dynamic_range=1e-6, odd_only=True):
    """Highest harmonic whose intensity is within ``dynamic_range`` of the strongest one."""
    sel = (q % 2 == 1) if odd_only else np.ones_like(q, dtype=bool)
    thr = dynamic_range * I[sel].max()
    above = q[sel & (I >= thr)]
    return int(above.max()) if len(above) else 0
