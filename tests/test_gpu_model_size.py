"""Models beyond the compact sizes: wide R shells (up to 256 placed support
points; the reference places any number, src/hamiltonian.cpp:8-39 and
src/wannier_tb.cpp:15-113) and 33-64 bands, against the oracle on small grids.

Kernels: shells with more than 8 distinct n2 values or more than 32 +-R
pairs of a 2-band model run on the element kernel (3); 8-band FP32 solves
whose staged blocks exceed the LDS take the lean row-separable form (9);
33-64 bands run on the 1024-lane element form (3).
"""
import numpy as np
import pytest

from conftest import make_param, rel_l2

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32_TOL = 1e-5


def _model(tmp_path, nb, nmax, seed):
    from hhgamd import tbgen
    path = str(tmp_path / f"syn{nb}_n{nmax}_tb.dat")
    tbgen.write_tb(path, *tbgen.synthetic(nb, seed=seed, nmax=nmax))
    return path


@pytest.mark.parametrize("nb,nmax,prec,kernel", [(4, 5, "fp64", 3), (2, 4, "fp64", 3), (8, 5, "fp64", 3),
                                                 (8, 3, "fp32", 9), (6, 7, "fp64", 3)])
def test_wide_support_shells_vs_oracle(tmp_path, oracle_lib, nb, nmax, prec, kernel):
    """(2 nmax + 1)^2 R vectors: 81 (9x9), 121 (11x11), 225 (15x15) support points."""
    from hhgamd import lib
    path = _model(tmp_path, nb, nmax, 70 + nb + nmax)
    p = dict(make_param(path, nr=(16, 16), tmax=3.0, tmax_field=2.5, intensity=2e12, pol=(0.6, 0.8, 0.0)),
             precision=prec)
    with lib.Solve(p) as s:
        kind = s.kernel()
        s.run()
        J = s.current()[1]
    _, Jref = oracle_lib.run(path, p, mode="perk")
    err = rel_l2(J, Jref)
    print(f"nb={nb} {(2 * nmax + 1) ** 2} R vectors {prec}: kernel {kind}, rel L2 vs oracle {err:.2e}")
    assert kind == kernel
    assert err <= (FP64_TOL if prec == "fp64" else FP32_TOL)


def test_support_beyond_the_limit_is_refused(tmp_path):
    from hhgamd import lib
    path = _model(tmp_path, 2, 8, 5)  # 17 x 17 = 289 R vectors > 256
    with pytest.raises(RuntimeError, match="support"):
        lib.Solve(make_param(path, nr=(40, 40), tmax=1.0))


@pytest.mark.parametrize("nb,nr,prec", [(40, (4, 6), "fp64"), (40, (4, 6), "fp32"), (64, (4, 4), "fp64")])
def test_more_than_32_bands_vs_oracle(tmp_path, oracle_lib, nb, nr, prec):
    from hhgamd import lib
    path = _model(tmp_path, nb, 1, 90 + nb)
    p = dict(make_param(path, nr=nr, tmax=0.5, tmax_field=0.45, intensity=2e12, pol=(0.6, 0.8, 0.0)),
             precision=prec)
    with lib.Solve(p) as s:
        kind, L = s.kernel(), s.launch()
        rho_dev = s.rho_k()
        s.run()
        J = s.current()[1]
    with lib.Solve(p, device=False) as s:
        rho_host = s.rho_k()
    _, Jref = oracle_lib.run(path, p, mode="perk")
    err = rel_l2(J, Jref)
    print(f"nb={nb} {nr} {prec}: kernel {kind} {L}, rel L2 vs oracle {err:.2e}, "
          f"device vs host initial state {np.abs(rho_dev - rho_host).max():.1e}")
    assert kind == 3 and L["b"] == 1  # the 1024-lane element form
    assert np.abs(rho_dev - rho_host).max() < 1e-12
    assert err <= (FP64_TOL if prec == "fp64" else FP32_TOL)


def test_more_than_64_bands_is_refused(tmp_path):
    from hhgamd import lib
    path = _model(tmp_path, 66, 0, 3)
    with pytest.raises(RuntimeError):
        lib.Solve(make_param(path, nr=(4, 4), tmax=0.2))
