"""The precision, batching and serving configurations at their full workloads.

C3  seeded 8-band Wannier model (5x5 R shell), 120x120 k-grid, 4096 RK4 steps,
    I = 5e11, in FP64 and with FP32 stages: final rho_k at 64 sampled
    k-points against the oracle's per-k propagation of the same k-points, the
    last current sample against the window formula of src/observable.cpp:16-30
    evaluated on the device's final states, FP32 against FP64.
C4  64 seeded instances (pol angle, intensity, wavelength, nb in {2, 4, 8},
    uneven 5-smooth grids in [60, 200]) in one batched run with FP32 stages:
    every member against its solo run, against its FP64 solve, and the two
    smallest 2-band members against the oracle.
C5  a Poisson stream (nb in {2, 4, 8, 16}, grids in [48, 240], tmax 60-150)
    through the server at 20 requests/s: every served result against the same
    request served alone (bitwise), run by itself (<= 1e-13: a served solve
    takes the launch rules of batch members) and against its FP64 solve.

Gates of the mixed path (FP32 stages): J(t) relative L2 <= 1e-5 against FP64
and every harmonic intensity within 1 % of FP64 up to the cutoff.
"""
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

J_TOL = 1e-5
HARMONIC_TOL = 0.01
FP64_TOL = 1e-10


def gates(t, J, Jref, lam):
    from hhgamd import spectrum
    err = rel_l2(J, Jref)
    worst = 0.0
    for a in range(3):
        if np.abs(Jref[:, a]).max() == 0:
            continue
        q, I = spectrum.harmonic_intensities(t, Jref[:, a], lam)
        _, I2 = spectrum.harmonic_intensities(t, J[:, a], lam)
        sel = q <= spectrum.cutoff_order(q, I)
        worst = max(worst, float(np.max(np.abs(I2[sel] / I[sel] - 1.0))))
    return err, worst


def window_current_last(path, prm, rho, A_last):
    """J_a[nt-1] from the final states rho (N, nb, nb): (1/N) sum_k sum_{n in the
    4x4 window} sum_ij v_a(n)_ij conj(rho_k,ij) e^{-i(k - A).R(n)}, v from the oracle."""
    import oracle
    lat, _, _, _ = oracle.model(path)
    v = oracle.velocity_window(path, prm)
    N1, N2 = prm["nr1"], prm["nr2"]
    m1, m2 = np.divmod(np.arange(N1 * N2), N2)
    rc = np.conj(rho).reshape(N1 * N2, -1)
    J = np.zeros(3, dtype=complex)
    for i1, n1 in enumerate(range(-2, 2)):
        for i2, n2 in enumerate(range(-2, 2)):
            R = (n1 * lat[0] + n2 * lat[1]) / 0.5291772
            ph = np.exp(-2j * np.pi * (m1 * n1 / N1 + m2 * n2 / N2)) * np.exp(1j * A_last @ R)
            for a in range(3):
                J[a] += ph @ (rc @ v[a, i1, i2].reshape(-1))
    return J / (N1 * N2)


def sample_k(N1, N2, n=64, seed=3):
    rng = np.random.default_rng(seed)
    ks = set([0, N1 * N2 - 1, N2 - 1, (N1 - 1) * N2, (N1 // 2) * N2 + N2 // 2])
    while len(ks) < n:
        ks.add(int(rng.integers(N1 * N2)))
    return np.array(sorted(ks), dtype=np.int32)


@pytest.fixture(scope="module")
def workdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("configs"))


@pytest.fixture(scope="module")
def c3_runs(workdir):
    """C3 solved on the device in FP64 and with FP32 stages; the oracle's
    final rho_k of 64 sampled k-points."""
    import oracle
    from hhgamd import lib, workloads
    prm = workloads.c3(workdir, precision="fp64")
    out = {"prm": prm}
    for prec in ("fp64", "fp32"):
        with lib.Solve(dict(prm, precision=prec)) as s:
            assert s.nb == 8 and s.N == 14400 and s.nt == 4096
            s.run()
            t, J = s.current()
            _, _, A = s.fields()
            out[prec] = {"t": t, "J": J, "rho": s.rho_k(), "A_last": A[:, -1], "kernel": s.kernel()}
    ks = sample_k(prm["nr1"], prm["nr2"])
    out["ks"] = ks
    out["rho_oracle"] = oracle.propagate_k(prm["path_tb"], prm, ks)
    return out


def test_c3_fp64_sampled_k_vs_oracle(c3_runs):
    r = c3_runs["fp64"]
    assert r["kernel"] == 6  # the row-separable kernel
    d = np.abs(r["rho"][c3_runs["ks"]] - c3_runs["rho_oracle"]).max()
    print(f"C3 FP64: max |rho_gpu - rho_oracle| over {len(c3_runs['ks'])} k-points = {d:.2e}")
    assert d <= FP64_TOL


def test_c3_fp64_last_current_is_the_window_sum_of_the_final_states(c3_runs):
    r = c3_runs["fp64"]
    Jw = window_current_last(c3_runs["prm"]["path_tb"], c3_runs["prm"], r["rho"], r["A_last"])
    err = np.abs(Jw - r["J"][-1]).max() / np.abs(r["J"][-1]).max()
    print(f"C3 FP64: J[nt-1] vs window sum of the device's final states: {err:.2e}")
    assert err <= 1e-12


def test_c3_fp32_gates_and_sampled_k(c3_runs):
    r32, r64 = c3_runs["fp32"], c3_runs["fp64"]
    err, worst = gates(r64["t"], r32["J"], r64["J"], c3_runs["prm"]["lambda"])
    drho = np.abs(r32["rho"][c3_runs["ks"]] - c3_runs["rho_oracle"]).max()
    print(f"C3 FP32: J rel L2 vs FP64 {err:.2e}, worst harmonic {worst:.2e}, "
          f"max |rho_fp32 - rho_oracle| {drho:.2e}")
    assert err <= J_TOL and worst <= HARMONIC_TOL
    assert drho <= 1e-5


@pytest.fixture(scope="module")
def c4_runs(workdir):
    from hhgamd import lib, workloads
    params = workloads.c4(workdir, n=64, precision="fp32")
    solves = [lib.Solve(p) for p in params]
    try:
        with lib.Batch(solves) as b:
            b.run()
        batched = [s.current()[1] for s in solves]
        t_all = [s.current()[0] for s in solves]
    finally:
        for s in solves:
            s.close()
    return params, t_all, batched


def test_c4_batch_members_equal_solo_runs(c4_runs):
    from hhgamd import lib
    params, _, batched = c4_runs
    worst = 0.0
    for p, Jb in zip(params, batched):
        with lib.Solve(p) as s:
            s.run()
            worst = max(worst, rel_l2(Jb, s.current()[1]))
    print(f"C4: 64 batched members vs solo runs, worst rel L2 = {worst:.1e}")
    # 5-8-band members keep one k-point per lane group in a batch; solo runs on
    # large grids may carry two, which reorders only the current's k-sum
    assert worst <= 1e-13


def test_c4_fp32_gates_vs_fp64(c4_runs):
    from hhgamd import lib
    params, t_all, batched = c4_runs
    errs, worsts = [], []
    for p, t, Jb in zip(params, t_all, batched):
        with lib.Solve(dict(p, precision="fp64")) as s:
            s.run()
            e, w = gates(t, Jb, s.current()[1], p["lambda"])
        errs.append(e)
        worsts.append(w)
    print(f"C4 FP32 batch vs FP64: max J rel L2 {max(errs):.2e}, worst harmonic {max(worsts):.2e}")
    assert max(errs) <= J_TOL and max(worsts) <= HARMONIC_TOL


def test_c4_small_two_band_members_vs_oracle(c4_runs, oracle_lib):
    params, t_all, batched = c4_runs
    two = sorted((i for i, p in enumerate(params) if oracle_lib.model_dims(p["path_tb"])[0] == 2),
                 key=lambda i: params[i]["nr1"] * params[i]["nr2"])[:2]
    assert two
    for i in two:
        p = params[i]
        _, Jref = oracle_lib.run(p["path_tb"], p, mode="perk")
        e, w = gates(t_all[i], batched[i], Jref, p["lambda"])
        print(f"C4 member {i} ({p['nr1']}x{p['nr2']}, 2 bands) FP32 batch vs oracle: J {e:.2e}, harmonic {w:.2e}")
        assert e <= J_TOL and w <= HARMONIC_TOL


def test_c5_poisson_stream_mix_equals_alone_and_fp64_gates(workdir):
    from hhgamd import lib, serve, workloads
    arrivals = workloads.c5(workdir, 24, rate=20.0, seed=7)
    with serve.Server(hbm_budget_gb=64.0, max_inflight=8, host_threads=8) as srv:
        recs, summ = serve.run_stream(srv, arrivals, keep_results=True)
    print(f"C5 stream: {summ}")
    assert summ["completed"] == len(arrivals) and summ["peak_inflight"] > 1
    errs, worsts = [], []
    with serve.Server(hbm_budget_gb=64.0, max_inflight=1, host_threads=1) as one:  # each request alone
        alone = [one.result(one.submit(p))[1] for _, p in arrivals]
    for rec, (_, p), Ja in zip(recs, arrivals, alone):
        assert np.array_equal(rec["J"], Ja), "served result differs from the same request served alone"
        with lib.Solve(p) as s:  # the same solve by itself (launch rules of a single solve)
            s.run()
            assert rel_l2(rec["J"], s.current()[1]) <= 1e-13
        with lib.Solve(dict(p, precision="fp64")) as s:
            s.run()
            e, w = gates(rec["t"], rec["J"], s.current()[1], p["lambda"])
        errs.append(e)
        worsts.append(w)
    print(f"C5 served FP32 vs FP64: max J rel L2 {max(errs):.2e}, worst harmonic {max(worsts):.2e}")
    assert max(errs) <= J_TOL and max(worsts) <= HARMONIC_TOL
