"""Batched propagation (hhg_batch_*): each member of a mixed batch gives the
same current as its solve run alone, and the oracle's current."""
import numpy as np
import pytest

from conftest import HALDANE, HMCASE0, make_param, rel_l2

pytestmark = pytest.mark.gpu

SAME_TOL = 1e-13     # batch member vs the same solve alone (identical arithmetic)
FP64_TOL = 1e-10     # vs the oracle


def _mixed_params(tmp_path):
    from hhgamd import tbgen, workloads
    files = {}
    for nb in (2, 3, 4, 8):
        files[nb] = str(tmp_path / f"syn{nb}_tb.dat")
        tbgen.write_tb(files[nb], *tbgen.synthetic(nb, seed=20 + nb))
    short = dict(tmax=8.0, tmax_field=7.0, intensity=2e12)
    # Haldane model shifted by -1 eV on site: two filled bands near the K points, one elsewhere
    lat, R, H, r = tbgen.haldane()
    H[[tuple(x[:2]) for x in R].index((0, 0))] += -1.0 * np.eye(2)
    doped = str(tmp_path / "haldane_doped_tb.dat")
    tbgen.write_tb(doped, lat, R, H, r)
    return [
        make_param(HALDANE, nr=(16, 20), pol=(0.6, 0.8, 0.0), **short),
        make_param(HMCASE0, nr=(12, 12), pol=(0.0, 1.0, 0.0), **short),
        make_param(HALDANE, nr=(10, 14), pol=(1.0, 0.0, 0.0), tmax=5.0, tmax_field=4.0, intensity=1e12),
        make_param(files[2], nr=(8, 10), pol=(0.6, 0.8, 0.0), **short),
        make_param(files[3], nr=(6, 8), pol=(0.6, 0.8, 0.0), **short),
        make_param(files[4], nr=(8, 8), pol=(0.6, 0.8, 0.0), **short),
        make_param(files[8], nr=(6, 6), pol=(0.8, 0.6, 0.0), tmax=4.0, tmax_field=3.0, intensity=2e12),
        dict(make_param(files[4], nr=(6, 8), pol=(0.6, 0.8, 0.0), **short), precision="fp32"),
        make_param(doped, nr=(16, 20), pol=(0.6, 0.8, 0.0), **short),
        # a filled lower band at every k-point: streamed pairs in Bloch-vector form
        make_param(workloads.synthetic_file(str(tmp_path), 2), nr=(8, 12), pol=(0.6, 0.8, 0.0), **short),
    ]


def test_batch_members_equal_single_runs(tmp_path):
    from hhgamd import lib
    params = _mixed_params(tmp_path)
    alone = []
    for p in params:
        with lib.Solve(p) as s:
            s.run()
            alone.append(s.current()[1])
    solves = [lib.Solve(p) for p in params]
    kinds = sorted({s.kernel() for s in solves})
    # register pairs (4-component and Bloch-vector), streamed pairs (4-component and Bloch-vector) and
    # row-separable kernels in one batch; alone, the register Bloch-vector solves take the two-wave form (10), in the batch
    # the one-wave form (8)
    assert kinds == [2, 4, 6, 10, 11]
    assert sorted({s.batch_launch()["kernel"] for s in solves}) == [2, 4, 6, 8, 11]
    with lib.Batch(solves) as b:
        b.run()
    for p, s, Ja in zip(params, solves, alone):
        Jb = s.current()[1]
        err = rel_l2(Jb, Ja)
        print(f"nr={p['nr1']}x{p['nr2']} kernel={s.kernel()} batch vs alone rel L2 = {err:.1e} "
              f"bitwise={np.array_equal(Jb, Ja)}")
        assert err <= SAME_TOL
        if s.kernel() == 10:  # the two forms run the same operations in the same order
            assert np.array_equal(Jb, Ja)
        s.close()


def test_batch_vs_oracle_and_continuation(tmp_path, oracle_lib):
    from hhgamd import lib
    params = _mixed_params(tmp_path)[:3]
    solves = [lib.Solve(p) for p in params]
    with lib.Batch(solves) as b:
        b.run()
        rho1 = [s.rho_k() for s in solves]
        b.run()  # continues from the evolved states; the first nt samples stay
    for p, s, r1 in zip(params, solves, rho1):
        _, J = s.current()
        _, Jref = oracle_lib.run(p["path_tb"], p, mode="perk")
        assert rel_l2(J, Jref) <= FP64_TOL
        assert np.abs(s.rho_k() - r1).max() > 0
        s.close()


def test_batch_with_element_kernel_members(tmp_path, oracle_lib):
    """Members on the element kernel (20 and 32 bands: one k-point per
    256-lane workgroup; a 6-band model forced to the generic kernels) batched
    with a row-separable one: each equals its solo run and the oracle."""
    from hhgamd import lib, tbgen
    files = {}
    for nb in (6, 20, 32):
        files[nb] = str(tmp_path / f"big{nb}_tb.dat")
        tbgen.write_tb(files[nb], *tbgen.synthetic(nb, seed=60 + nb))
    short = dict(tmax=1.5, tmax_field=1.2, intensity=2e12, pol=(0.6, 0.8, 0.0))
    params = [make_param(files[20], nr=(6, 6), **short), make_param(files[32], nr=(4, 6), **short),
              dict(make_param(files[6], nr=(6, 8), **short), kernel="generic"),
              make_param(files[6], nr=(8, 6), **short)]
    alone = []
    for p in params:
        with lib.Solve(p) as s:
            s.run()
            alone.append(s.current()[1])
    solves = [lib.Solve(p) for p in params]
    assert [s.kernel() for s in solves] == [3, 3, 3, 6]
    with lib.Batch(solves) as b:
        b.run()
    for p, s, Ja in zip(params, solves, alone):
        Jb = s.current()[1]
        _, Jref = oracle_lib.run(p["path_tb"], p, mode="perk")
        print(f"kernel={s.kernel()} batch vs alone {rel_l2(Jb, Ja):.1e}, vs oracle {rel_l2(Jb, Jref):.1e}")
        assert rel_l2(Jb, Ja) <= SAME_TOL
        assert rel_l2(Jb, Jref) <= FP64_TOL
        s.close()


def test_member_changes_after_adding(tmp_path):
    """A batch resolves its members at every run: a member restarted after it
    was added starts again from equilibrium, a member switched to FP32 stages
    runs them, and a deleted member leaves the batch."""
    from hhgamd import lib
    params = _mixed_params(tmp_path)
    pa, pb, pc = params[0], params[5], params[1]   # Haldane 16x20, 4-band 8x8, hmcase0 12x12
    with lib.Solve(pa) as s:
        s.run()
        Ja = s.current()[1]
    with lib.Solve(pb) as s:
        s.run()
        s.set_precision("fp32")
        s.run()
        Jb = s.current_run(-1)[1]
    A, B, C = lib.Solve(pa), lib.Solve(pb), lib.Solve(pc)
    b = lib.Batch([A, B, C])
    try:
        b.run()
        A.restart()
        B.set_precision("fp32")
        C.close()
        assert b.lib.hhg_batch_size(b.ptr) == 2
        b.run()
        ea, eb = rel_l2(A.current()[1], Ja), rel_l2(B.current_run(-1)[1], Jb)
        print(f"restarted member vs a fresh solo run {ea:.1e}; FP32-switched member vs the same switch solo {eb:.1e}")
        assert A.current()[1].shape == Ja.shape and ea <= SAME_TOL
        assert eb <= SAME_TOL
    finally:
        b.close()
        A.close()
        B.close()


@pytest.mark.parametrize("model,nr,nt,prec", [
    ("haldane", (160, 200), None, "fp64"),   # 500 groups of 64 k-points, 4096 steps: the largest two-wave grids
    ("haldane", (30, 30), None, "fp64"), ("haldane", (30, 30), None, "fp32"),
    ("haldane", (18, 22), 2, "fp64"), ("haldane", (18, 22), 3, "fp64"), ("haldane", (18, 22), 4, "fp64"),
    ("haldane", (10, 6), 16, "fp64"), ("haldane", (10, 6), 17, "fp64"), ("haldane", (10, 6), 33, "fp64"),
    ("hmcase0", (12, 12), 40, "fp64"), ("hmcase0", (14, 10), 47, "fp64"), ("hmcase0", (12, 12), 40, "fp32"),
])
def test_two_wave_bloch_form_equals_one_wave_form(model, nr, nt, prec):
    """A Bloch-vector solve alone on a grid of at most two groups of 64
    k-points per CU (512 groups on 256 CUs) runs the two-wave form (kind 10);
    as the member of a batch it runs the one-wave form (kind 8). Both form every value
    by the same operations in the same order: J(t) and the final state are
    bitwise equal over two consecutive runs, for ragged last groups of 64
    k-points, step counts around the 16-step current chunks and the two state
    slots, with rbc terms (hmcase0 at oblique polarisation) and FP32 stages."""
    from hhgamd import lib
    path = {"haldane": HALDANE, "hmcase0": HMCASE0}[model]
    kw = {} if nt is None else dict(tmax=0.02197 * nt + 1e-9, dt=0.02197, tmax_field=0.02197 * nt * 0.9,
                                    intensity=4e12)
    prm = dict(make_param(path, nr=nr, pol=(0.6, 0.8, 0.0) if model == "hmcase0" else (0.0, 1.0, 0.0), **kw),
               precision=prec)
    with lib.Solve(prm) as a, lib.Solve(prm) as b:
        la, lb = a.launch(), b.batch_launch()
        print(f"{model} {nr} nt={a.nt}: alone {la}, batched {lb}")
        assert la["kernel"] == 10 and lb["kernel"] == 8 and (la["a"], la["b"], la["c"]) == (lb["a"], lb["b"], lb["c"])
        if nt is not None:
            assert a.nt == nt
        with lib.Batch([b]) as bt:
            for _ in range(2):
                a.run()
                bt.run()
        steps = a.nt
        Ja = np.stack([a.current_run(0)[1], a.current_run(1)[1]])
        Jb = np.stack([b.current_run(0)[1], b.current_run(1)[1]])
        ra, rb = a.rho_k(), b.rho_k()
    assert Ja.shape == Jb.shape == (2, steps, 3)
    assert np.array_equal(Ja, Jb) and np.array_equal(ra, rb)


def test_bloch_form_by_grid_size():
    """The 200 x 200 README-scale grid (625 groups) keeps the one-wave form
    alone; 160 x 200 (500 groups) takes the two-wave form."""
    from hhgamd import lib
    with lib.Solve(make_param(HALDANE, nr=200)) as s:
        assert s.kernel() == 8
    with lib.Solve(make_param(HALDANE, nr=(160, 200))) as s:
        assert s.kernel() == 10
