"""Autotuner: the tuned launch configuration changes only the work layout,
never the state it starts from, and its results match the default launch
and the oracle."""
import numpy as np
import pytest

from conftest import HALDANE, make_param, rel_l2

pytestmark = pytest.mark.gpu


def _syn(tmp_path, nb, seed):
    from hhgamd import tbgen
    path = str(tmp_path / f"syn{nb}_tb.dat")
    tbgen.write_tb(path, *tbgen.synthetic(nb, seed=seed))
    return path


@pytest.mark.parametrize("nb,nr", [(4, (10, 40)), (8, (8, 30))])
def test_autotuned_separable_matches_default_and_oracle(tmp_path, oracle_lib, nb, nr):
    from hhgamd import lib
    path = _syn(tmp_path, nb, 60 + nb)
    p = make_param(path, nr=nr, tmax=6.0, tmax_field=5.0, intensity=2e12, pol=(0.6, 0.8, 0.0))
    with lib.Solve(p) as s:
        default = s.launch()
        s.run()
        J0 = s.current()[1]
    with lib.Solve(p) as s:
        rho0 = s.rho_k()
        kind = s.autotune(steps=16)
        assert np.array_equal(s.rho_k(), rho0)  # tuning runs leave the state alone
        tuned = s.launch()
        s.run()
        J1 = s.current()[1]
    print(f"nb={nb}: default {default}, tuned {tuned}")
    assert kind == tuned["kernel"] == 6 and 1 <= tuned["b"] <= 8
    assert rel_l2(J1, J0) <= 1e-12
    _, Jref = oracle_lib.run(path, p, mode="perk")
    assert rel_l2(J1, Jref) <= 1e-10


def test_autotuned_pair_kernels_match_oracle(oracle_lib):
    from hhgamd import lib
    p = make_param(HALDANE, nr=(16, 20), tmax=10.0, tmax_field=8.0, pol=(0.6, 0.8, 0.0))
    with lib.Solve(p) as s:
        kind = s.autotune(steps=16)
        s.run()
        J = s.current()[1]
    assert kind in (1, 2, 4, 8, 10)
    _, Jref = oracle_lib.run(HALDANE, p, mode="perk")
    assert rel_l2(J, Jref) <= 1e-10


def test_process_autotune_batch_and_cache(tmp_path):
    from hhgamd import lib
    path = _syn(tmp_path, 4, 71)
    p = dict(make_param(path, nr=(8, 24), tmax=4.0, tmax_field=3.0, pol=(0.6, 0.8, 0.0)), precision="fp32")
    alone = []
    for _ in range(2):
        with lib.Solve(p) as s:
            s.run()
            alone.append(s.current()[1])
    lib.set_autotune(True)
    try:
        solves = [lib.Solve(p) for _ in range(2)]
        with lib.Batch(solves) as b:
            b.run()
        keys = [s.launch() for s in solves]
        assert keys[0] == keys[1]  # same shape: one tuning, cached
        for s, Ja in zip(solves, alone):
            assert rel_l2(s.current()[1], Ja) <= 1e-12
            s.close()
    finally:
        lib.set_autotune(False)
