"""Mixed precision (FP32 RK4 stages, FP64 state/initial state/current) against
the FP64 path on the GPU.

Tolerances (targets of the mixed path): J(t) relative L2 <= 1e-5 against FP64
and every harmonic intensity within 1 % of FP64 up to the cutoff, where the
cutoff is the highest odd harmonic within 1e-6 of the strongest one.
"""
import numpy as np
import pytest

from conftest import HALDANE, HMCASE0, make_param, rel_l2

pytestmark = pytest.mark.gpu

J_TOL = 1e-5
HARMONIC_TOL = 0.01


def run(prm):
    from hhgamd import lib
    with lib.Solve(prm) as s:
        s.run()
        return s.current()


def check(prm, lam=3000.0):
    from hhgamd import spectrum
    t, J64 = run(dict(prm, precision="fp64"))
    _, J32 = run(dict(prm, precision="fp32"))
    err = rel_l2(J32, J64)
    worst = 0.0
    for a in range(3):
        if np.abs(J64[:, a]).max() == 0:
            continue
        q, I64 = spectrum.harmonic_intensities(t, J64[:, a], lam)
        _, I32 = spectrum.harmonic_intensities(t, J32[:, a], lam)
        qc = spectrum.cutoff_order(q, I64)
        sel = q <= qc
        worst = max(worst, float(np.max(np.abs(I32[sel] / I64[sel] - 1.0))))
    print(f"J rel L2 fp32 vs fp64 = {err:.3e}, worst harmonic deviation up to cutoff = {worst:.3e}")
    return err, worst


def test_fp32_c2_haldane_200():
    err, worst = check(make_param(HALDANE, nr=200))
    assert err <= J_TOL and worst <= HARMONIC_TOL


def test_fp32_hmcase0_readme():
    err, worst = check(make_param(HMCASE0, nr=100))
    assert err <= J_TOL and worst <= HARMONIC_TOL


@pytest.mark.parametrize("nb,nr", [(2, 24), (3, 24), (4, 24), (8, 24), (16, 24), (20, 12)])
def test_fp32_synthetic(tmp_path, nb, nr):
    from hhgamd import tbgen
    path = str(tmp_path / f"syn{nb}_tb.dat")
    tbgen.write_tb(path, *tbgen.synthetic(nb, seed=11 + nb))
    err, worst = check(make_param(path, nr=nr, intensity=5e11, pol=(0.6, 0.8, 0.0)))
    assert err <= J_TOL and worst <= HARMONIC_TOL


def test_fp32_synthetic_16_bands_7x7_shell(tmp_path):
    """16 bands over a 7x7 R shell: 7 n2 values, the first without a velocity
    block (the velocity tables hold the other values by rank)."""
    from hhgamd import lib, tbgen
    path = str(tmp_path / "syn16n3_tb.dat")
    tbgen.write_tb(path, *tbgen.synthetic(16, seed=31, nmax=3))
    prm = make_param(path, nr=16, intensity=5e11, pol=(0.6, 0.8, 0.0))
    with lib.Solve(dict(prm, precision="fp32")) as s:
        assert s.kernel() == 6 and s.launch()["c"] == 8
    err, worst = check(prm)
    assert err <= J_TOL and worst <= HARMONIC_TOL


@pytest.mark.parametrize("nb", [2, 3, 4, 8, 16])
def test_fp32_state_stays_a_hermitian_projector(tmp_path, nb):
    """After a full FP32-stage run each rho_k is exactly Hermitian (both the
    M - M^dagger form and, for 3-4 bands, the matrix-core HX - XH form build
    exactly Hermitian stage derivatives), keeps its initial trace to FP32
    accuracy, stays as close to a projector as the FP64 path's RK4 state, and
    agrees with that state. A commutator whose rounding breaks the symmetry
    (Hermiticity ~1e-7, trace drift ~3e-5 here) fails the first two."""
    from hhgamd import lib, tbgen
    path = str(tmp_path / f"syn{nb}_tb.dat")
    tbgen.write_tb(path, *tbgen.synthetic(nb, seed=70 + nb))
    prm = make_param(path, nr=16 if nb > 8 else 24, intensity=5e11, pol=(0.6, 0.8, 0.0))
    states = {}
    for prec in ("fp64", "fp32"):
        with lib.Solve(dict(prm, precision=prec)) as s:
            rho0 = s.rho_k()
            s.run()
            states[prec] = s.rho_k()
    idem = {k: np.abs(np.einsum("kij,kjl->kil", r, r) - r).max() for k, r in states.items()}
    rho = states["fp32"]
    herm = np.abs(rho - np.conj(np.transpose(rho, (0, 2, 1)))).max()
    dtr = np.abs(np.trace(rho, axis1=1, axis2=2) - np.trace(rho0, axis1=1, axis2=2)).max()
    d64 = np.abs(rho - states["fp64"]).max()
    print(f"nb={nb}: |rho - rho^H| {herm:.1e}, |dTr| {dtr:.1e}, |rho^2 - rho| {idem['fp32']:.1e} "
          f"(FP64 {idem['fp64']:.1e}), |rho32 - rho64| {d64:.1e}")
    assert herm < 1e-12 and dtr < 1e-5 and idem["fp32"] < idem["fp64"] + 1e-5 and d64 < 1e-5
    # the FP64 path: exactly Hermitian, trace to rounding
    r64 = states["fp64"]
    assert np.abs(r64 - np.conj(np.transpose(r64, (0, 2, 1)))).max() < 1e-14
    assert np.abs(np.trace(r64, axis1=1, axis2=2) - np.trace(rho0, axis1=1, axis2=2)).max() < 1e-11
