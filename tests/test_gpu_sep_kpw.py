"""The row-separable kernel with several k-points per lane group.

When an 8- or 16-band grid has more workgroups than the device has CUs, each
lane group carries 2 (8 bands) or 2-4 (16 bands) k-points of its row through a
step under one per-row table build (SWESim_get_kpoints_per_wave). The k-points
of a group sit W groups apart in the row, and the last tile of a row may be
partly past its end. These runs must match the oracle, and batched they must
equal their runs alone.
"""
import numpy as np
import pytest

from conftest import make_param, rel_l2

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10   # vs the oracle
FP32_TOL = 1e-5    # J of FP32 stages vs the FP64 oracle
SAME_TOL = 1e-13   # batch member vs the same solve alone

# (bands, grid, precision) whose workgroups outnumber the CUs. 16 bands (8
# waves per workgroup): rows of 22 carry 4 k-points per wave in tiles of 8 x 4
# (the second partly past the row's end), rows of 20 carry 2 in tiles of 7 x 2,
# rows of 40 carry 2; 12 bands padded to 16 likewise; 8 and 6 (padded to 8)
# bands: 5 waves of 4 k-points, rows of 36
CASES = [(16, (200, 22), "fp64"), (16, (100, 20), "fp64"), (16, (80, 40), "fp64"), (16, (80, 40), "fp32"),
         (12, (200, 22), "fp64"), (12, (80, 40), "fp32"),
         (8, (160, 36), "fp64"), (8, (160, 36), "fp32"), (6, (160, 36), "fp64")]


def _model(tmp_path, nb):
    from hhgamd import tbgen
    path = str(tmp_path / f"syn{nb}_tb.dat")
    tbgen.write_tb(path, *tbgen.synthetic(nb, seed=90 + nb))
    return path


def _param(path, nr, prec):
    return dict(make_param(path, nr=nr, tmax=1.0, tmax_field=0.8, intensity=2e12, pol=(0.6, 0.8, 0.0)),
                precision=prec)


@pytest.mark.parametrize("nb,nr,prec", CASES)
def test_kpoints_per_wave_vs_oracle(tmp_path, oracle_lib, nb, nr, prec):
    from hhgamd import lib
    path = _model(tmp_path, nb)
    p = _param(path, nr, prec)
    with lib.Solve(p) as s:
        kind, kpw = s.kernel(), s.kpoints_per_wave()
        s.run()
        J = s.current()[1]
        rho = s.rho_k()
    _, Jref = oracle_lib.run(path, p, mode="perk")
    err = rel_l2(J, Jref)
    print(f"nb={nb} nr={nr} {prec}: kernel {kind}, {kpw} k-points per lane group, rel L2 vs oracle {err:.2e}")
    assert kind == 6 and kpw in ((2, 4) if nb > 8 else (2,))
    assert err <= (FP64_TOL if prec == "fp64" else FP32_TOL)
    # every k-point of the grid, partial tiles included, was propagated: rho
    # stays Hermitian with its trace, and the field moved it off the start
    with lib.Solve(p, device=False) as s0:
        rho0 = s0.rho_k()
    assert np.abs(rho - np.conj(np.swapaxes(rho, 1, 2))).max() < 1e-12
    tr_tol = 1e-10 if prec == "fp64" else 1e-6  # FP32 stage increments round the trace at ~1e-7
    assert np.abs(np.trace(rho, axis1=1, axis2=2) - np.trace(rho0, axis1=1, axis2=2)).max() < tr_tol
    assert np.all(np.abs(rho - rho0).reshape(rho.shape[0], -1).max(axis=1) > 0)


def test_kpoints_per_wave_in_a_batch(tmp_path):
    from hhgamd import lib
    params = [_param(_model(tmp_path, nb), nr, prec) for nb, nr, prec in CASES]
    alone = []
    for p in params:
        with lib.Solve(p) as s:
            s.run()
            alone.append(s.current()[1])
    solves = [lib.Solve(p) for p in params]
    with lib.Batch(solves) as b:
        b.run()
    for p, s, Ja in zip(params, solves, alone):
        Jb = s.current()[1]
        print(f"nr=({p['nr1']}, {p['nr2']}) {p['precision']}: {s.kpoints_per_wave()} k-points per wave, "
              f"batch vs alone {rel_l2(Jb, Ja):.1e}")
        assert s.kpoints_per_wave() in (2, 4)
        assert rel_l2(Jb, Ja) <= SAME_TOL
        s.close()


def test_lean_form_of_the_8_band_fp32_kernel(tmp_path, oracle_lib):
    """FP32 solves of 5-8 bands whose lane groups carry one k-point take the
    lean form (at most 168 VGPRs, the blocks rounded to FP32 as the staged
    form holds them) with workgroups of up to 12 waves: rows of 40 k-points in
    one 10-wave workgroup. Against the oracle, and batched against alone (a
    batched member takes the lean form too)."""
    from hhgamd import lib
    params = [_param(_model(tmp_path, nb), (24, 40), "fp32") for nb in (8, 6)]
    alone = []
    for p in params:
        with lib.Solve(p) as s:
            L, kpw = s.launch(), s.kpoints_per_wave()
            s.run()
            J = s.current()[1]
        _, Jref = oracle_lib.run(p["path_tb"], p, mode="perk")
        err = rel_l2(J, Jref)
        print(f"nb={s.nb} 24x40 fp32: waves per workgroup {L['b']}, {kpw} k-point per lane group, "
              f"rel L2 vs oracle {err:.2e}")
        assert L["kernel"] == 9 and L["a"] == 8 and L["b"] == 10 and kpw == 1
        assert err <= FP32_TOL
        alone.append(J)
    solves = [lib.Solve(p) for p in params]
    with lib.Batch(solves) as b:
        b.run()
    for s, Ja in zip(solves, alone):
        assert rel_l2(s.current()[1], Ja) <= SAME_TOL
        s.close()


def test_batched_5_to_8_bands_carry_two_kpoints_on_long_rows(tmp_path, oracle_lib):
    """Batched 5-8-band members whose rows hold two tiles' worth carry two
    k-points per lane group (staged FP64 form and lean FP32 form, last tiles
    partly past the row end): each member matches its solo run to the
    summation order of its current and the FP64 one matches the oracle."""
    from hhgamd import lib
    cases = [(8, (40, 80), "fp64"), (6, (40, 90), "fp64"), (8, (40, 110), "fp32"), (6, (40, 100), "fp32")]
    params = [_param(_model(tmp_path, nb), nr, prec) for nb, nr, prec in cases]
    alone = []
    for p in params:
        with lib.Solve(p) as s:
            s.run()
            alone.append(s.current()[1])
    solves = [lib.Solve(p) for p in params]
    for (nb, nr, prec), s in zip(cases, solves):
        L = s.batch_launch()
        assert L["kernel"] == (9 if prec == "fp32" else 6) and L["kpw"] == 2, (nb, nr, prec, L)
    with lib.Batch(solves) as b:
        b.run()
    for (nb, nr, prec), p, s, Ja in zip(cases, params, solves, alone):
        Jb = s.current()[1]
        err = rel_l2(Jb, Ja)
        print(f"nb={nb} nr={nr} {prec}: batch vs alone {err:.1e}")
        assert err <= SAME_TOL
        if prec == "fp64":
            _, Jref = oracle_lib.run(p["path_tb"], p, mode="perk")
            assert rel_l2(Jb, Jref) <= FP64_TOL
        s.close()
