"""Polarisation scan: the multi-instance driver of scripts/pol_scan.py on one GPU.

The reference runs one solve per polarisation angle, one MPI rank per group of
angles (scripts/pol_scan.py:44-101): pol = (sin a, cos a, 0) for n_angles
values of a in [0, 2 pi], the spectrum fftshift(fft(ifftshift(gradient(J, dt))))
of each axis (scripts/pol_scan.py:23-42), and a record with the angles, the
harmonic axis freq / w0, w0 and the (n_angles, 3, nw) spectra
(scripts/pol_scan.py:86-101).

Here every angle is a solve of one batch (one persistent launch per kernel
family, ``mode="batch"``) or a request to the solve server (``mode="serve"``,
requests overlap on the device as they are admitted). Differences on purpose:
each angle's own spectra are stored (the reference's gather loop writes the last
angle's spectra into every row, scripts/pol_scan.py:90-92), and the record is
saved as .npz rather than a pickle.

Command line:
    python -m hhgamd.polscan --tb hmcase0_tb.dat --nr 200 --angles 15 --out scan.npz
"""
import argparse
import sys

import numpy as np

from . import spectrum


def angles(n_angles=15):
    """scripts/pol_scan.py:46: np.linspace(0, 2 pi, n_angles)."""
    return np.linspace(0, 2 * np.pi, n_angles)


def params_for(param, pol_angles):
    """One parameter dict per angle, pol = (sin a, cos a, 0) (scripts/pol_scan.py:24)."""
    out = []
    for a in pol_angles:
        p = dict(param)
        p["pol_vec"] = np.array([np.sin(a), np.cos(a), 0.0])
        p["phi_vec"] = np.asarray(param.get("phi_vec", np.zeros(3)), dtype=float)
        out.append(p)
    return out


def spectra(J, param):
    """(freq / w0, w0, spectra[3, nw]) of one solve's current J[nt, 3]."""
    dt, tmax, lam = param["dt"], param["tmax"], param["lambda"]
    freq_q, _ = spectrum.polscan_spectrum(J[:, 0], dt, tmax, lam)
    hhg = np.stack([spectrum.polscan_spectrum(J[:, a], dt, tmax, lam)[1] for a in range(3)])
    return freq_q, 2.0 * np.pi * 300 / lam, hhg


def run(param, n_angles=15, mode="batch", server=None):
    """Runs the scan and returns the record of scripts/pol_scan.py:95-98
    (pol_angles, freq in units of w0, omega0, data_sim[n_angles, 3, nw])."""
    from . import lib
    pol = angles(n_angles)
    plist = params_for(param, pol)
    if mode == "batch":
        solves = [lib.Solve(p) for p in plist]
        try:
            with lib.Batch(solves) as b:
                b.run()
            currents = [s.current()[1] for s in solves]
        finally:
            for s in solves:
                s.close()
    elif mode == "serve":
        from . import serve
        own = server is None
        srv = serve.Server() if own else server
        try:
            ids = [srv.submit(p) for p in plist]
            currents = []
            for i in ids:
                currents.append(srv.result(i)[1])
                srv.release(i)
        finally:
            if own:
                srv.close()
    else:
        raise ValueError("mode must be 'batch' or 'serve'")
    rows = [spectra(J, p) for J, p in zip(currents, plist)]
    return {"pol_angles": pol, "freq": rows[0][0], "omega0": rows[0][1],
            "data_sim": np.stack([r[2] for r in rows]).astype(np.complex128)}


def main(argv=None):
    ap = argparse.ArgumentParser(description="polarisation scan of HHG spectra (scripts/pol_scan.py)")
    ap.add_argument("--tb", required=True, help="Wannier _tb.dat file")
    ap.add_argument("--nr", type=int, nargs="+", default=[200], help="k-grid side(s)")
    ap.add_argument("--angles", type=int, default=15)
    ap.add_argument("--tmax", type=float, default=90.0)
    ap.add_argument("--dt", type=float, default=21.97e-3)
    ap.add_argument("--intensity", type=float, default=1e12)
    ap.add_argument("--lam", type=float, default=3000.0)
    ap.add_argument("--tmax-field", type=float, default=80.0)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--mode", default="batch", choices=["batch", "serve"])
    ap.add_argument("--out", default="pol_scan.npz")
    a = ap.parse_args(argv)
    n1, n2 = (a.nr[0], a.nr[0]) if len(a.nr) == 1 else (a.nr[0], a.nr[1])
    param = {"path_tb": a.tb, "nr1": n1, "nr2": n2, "nr3": 1, "tmax": a.tmax, "dt": a.dt,
             "intensity": a.intensity, "lambda": a.lam, "tmax_field": a.tmax_field,
             "pol_vec": np.array([0.0, 1.0, 0.0]), "phi_vec": np.zeros(3), "precision": a.precision}
    rec = run(param, a.angles, a.mode)
    np.savez(a.out, **rec)
    print(f"saved {a.out}: {len(rec['pol_angles'])} angles x 3 axes x {rec['data_sim'].shape[-1]} frequencies")
    return 0


if __name__ == "__main__":
    sys.exit(main())
