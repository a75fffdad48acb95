"""Writes the committed golden fixtures of tests/golden/.

Run from the repository root: ``python tests/golden/make_golden.py``.

* ``haldane_tb.dat`` -- the Haldane-like honeycomb model of the 30x30 and
  200x200 workloads (hhgamd.tbgen.haldane), in the reference's _tb.dat format.
* ``golden_<model>_<n>.npz`` -- t and J(t) (nt x 3 complex) from the oracle's
  FFT-form restatement of the reference's real-space time loop, after
  checking that the independent per-k restatement agrees to <= 1e-11.
  These are restatement-pinned values (the reference itself cannot be built
  here: FFTW3/LAPACKE/CBLAS are absent), used as regression anchors.
* ``fields_readme.npz`` -- t, E, A of the README pulse (float sinf/sqrtf
  semantics of src/efield.cpp:30-50).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "hhg-solids_amd"))

import oracle  # noqa: E402
from hhgamd import tbgen  # noqa: E402


def param(path, n, tmax=90.0):
    return {"path_tb": path, "nr1": n, "nr2": n, "nr3": 1, "tmax": tmax, "dt": 0.02197, "intensity": 1e12,
            "lambda": 3000.0, "tmax_field": 80.0, "pol_vec": [0.0, 1.0, 0.0], "phi_vec": [0.0, 0.0, 0.0]}


def main():
    oracle.build()
    hal = os.path.join(HERE, "haldane_tb.dat")
    tbgen.write_tb(hal, *tbgen.haldane(), comment="Haldane-like honeycomb: mass 0.3 eV, NN 2.8 eV, NNN 0.1 eV e^{+-i pi/2}")
    hm = os.path.join(HERE, "hmcase0_tb.dat")
    for name, path, n in [("hmcase0", hm, 10), ("hmcase0", hm, 30), ("haldane", hal, 30)]:
        prm = param(path, n)
        t, Jf = oracle.run(path, prm, mode="fft")
        _, Jk = oracle.run(path, prm, mode="perk")
        err = np.linalg.norm(Jk - Jf) / np.linalg.norm(Jf)
        print(f"{name} {n}x{n}: nt={len(t)} fft-vs-perk rel L2 = {err:.3e}")
        assert err < 1e-11
        np.savez_compressed(os.path.join(HERE, f"golden_{name}_{n}.npz"), t=t, J=Jf)
    t, E, A = oracle.fields(param(hm, 10))
    np.savez_compressed(os.path.join(HERE, "fields_readme.npz"), t=t, E=E, A=A)


if __name__ == "__main__":
    main()
