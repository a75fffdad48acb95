"""Synthetic workloads of the benchmark configurations, as parameter dicts.

Each solve is a dict with the keys of the reference's pySWE parameter
dictionary (path_tb, nr1, nr2, nr3, tmax, dt, intensity, lambda, tmax_field,
pol_vec, phi_vec) plus ``precision``. Models are generated into ``workdir``
with :mod:`hhgamd.tbgen` (fixed seeds), so every workload is reproducible.

  readme_pulse()   dt = 0.02197, tmax = 90 (4096 RK4 steps), lambda = 3000 nm,
                   I = 1e12 W/cm^2, sin^2 envelope to 80, pol along y
  c1 / c2          Haldane-like honeycomb, 30x30 / 200x200, FP64
  c3               seeded 8-band model (5x5 R shell), 120x120, I = 5e11, FP32 stages
  c4               64 instances: pol angle U[0, pi), I log-uniform 1e11..5e12,
                   lambda in {1600, 2400, 3200} nm, nb in {2, 4, 8}, sides drawn
                   from the even 5-smooth sizes in [60, 200]
  c5               Poisson request stream: nb in {2, 4, 8, 16}, sides from the
                   even 5-smooth sizes in [48, 240], tmax U[60, 150]
"""
import os

import numpy as np

from . import tbgen

DT = 0.02197


def smooth_even_sizes(lo, hi):
    """Even sizes n in [lo, hi] with no prime factor above 5."""
    out = []
    for n in range(lo + (lo % 2), hi + 1, 2):
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            out.append(n)
    return out


def readme_pulse(path_tb, nr, precision="fp64"):
    n1, n2 = (nr, nr) if isinstance(nr, int) else nr
    return {"path_tb": path_tb, "nr1": n1, "nr2": n2, "nr3": 1, "tmax": 90.0, "dt": DT, "intensity": 1e12,
            "lambda": 3000.0, "tmax_field": 80.0, "pol_vec": np.array([0.0, 1.0, 0.0]),
            "phi_vec": np.array([0.0, 0.0, 0.0]), "precision": precision}


def haldane_file(workdir):
    path = os.path.join(workdir, "haldane_tb.dat")
    if not os.path.exists(path):
        tbgen.write_tb(path, *tbgen.haldane(), comment="Haldane-like honeycomb (hhgamd.tbgen.haldane)")
    return path


def synthetic_file(workdir, nb, seed=None):
    seed = 1000 + nb if seed is None else seed
    path = os.path.join(workdir, f"synthetic_nb{nb}_s{seed}_tb.dat")
    if not os.path.exists(path):
        tbgen.write_tb(path, *tbgen.synthetic(nb, seed=seed), comment=f"seeded synthetic nb={nb} seed={seed}")
    return path


def c1(workdir):
    return readme_pulse(haldane_file(workdir), 30)


def c2(workdir, precision="fp64"):
    return readme_pulse(haldane_file(workdir), 200, precision)


def c3(workdir, precision="fp32", nb=8, nr=120):
    """Seeded nb-band model (5x5 R shell) on an nr x nr grid, I = 5e11, the
    README pulse otherwise (the 8-band 120 x 120 case by default)."""
    p = readme_pulse(synthetic_file(workdir, nb), nr, precision)
    p["intensity"] = 5e11
    return p


def c4(workdir, n=64, seed=4, precision="fp32", nbs=(2, 4, 8), lo=60, hi=200):
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = smooth_even_sizes(lo, hi)
    out = []
    for _ in range(n):
        th = rng.uniform(0.0, np.pi)
        nb = int(rng.choice(nbs))
        p = readme_pulse(synthetic_file(workdir, nb), (int(rng.choice(sizes)), int(rng.choice(sizes))), precision)
        p["pol_vec"] = np.array([np.sin(th), np.cos(th), 0.0])
        p["intensity"] = float(10 ** rng.uniform(11.0, np.log10(5e12)))
        p["lambda"] = float(rng.choice([1600.0, 2400.0, 3200.0]))
        out.append(p)
    return out


def c5(workdir, n, rate, seed=5, precision="fp32", nbs=(2, 4, 8, 16), lo=48, hi=240):
    """Returns [(arrival_time_s, param)] for a Poisson stream of ``n`` requests."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = smooth_even_sizes(lo, hi)
    t = 0.0
    out = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate)
        nb = int(rng.choice(nbs))
        p = readme_pulse(synthetic_file(workdir, nb), (int(rng.choice(sizes)), int(rng.choice(sizes))), precision)
        th = rng.uniform(0.0, np.pi)
        p["pol_vec"] = np.array([np.sin(th), np.cos(th), 0.0])
        p["intensity"] = float(10 ** rng.uniform(11.0, np.log10(5e12)))
        p["tmax"] = float(rng.uniform(60.0, 150.0))
        p["tmax_field"] = 0.9 * p["tmax"]
        out.append((t, p))
    return out


def ksteps(p):
    return p["nr1"] * p["nr2"] * int(p["tmax"] / p["dt"])
