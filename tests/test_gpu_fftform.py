"""The reference's real-space algorithm on the GPU (kernel variant 2): the
LDS-resident radix FFT against numpy, the FFT-form solve against the
oracle's FFT-form restatement and against the per-k kernels."""
import numpy as np
import pytest

from conftest import HALDANE, HMCASE0, make_param, rel_l2

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


@pytest.mark.parametrize("shape", [(1, 8, 10), (3, 12, 30), (2, 200, 240), (2, 14, 26), (1, 256, 256), (4, 60, 4),
                                   (2, 22, 66), (1, 143, 39), (3, 121, 169), (2, 49, 250), (1, 2, 3),
                                   (1, 400, 400), (2, 512, 270), (1, 288, 507), (2, 14, 484)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft2_matches_numpy(shape, sign):
    from hhgamd import lib
    rng = np.random.default_rng(sum(shape))
    a = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    got = lib.fft2(a, sign=sign)
    ref = np.fft.fft2(a) if sign < 0 else np.fft.ifft2(a) * (shape[1] * shape[2])
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"{shape} sign {sign}: max rel err {err:.2e}")
    assert err <= 1e-13


@pytest.mark.parametrize("shape", [(8, 200, 200), (3, 12, 30), (2, 256, 256), (2, 400, 400)])
def test_fft2_plan_on_device_memory(shape):
    """hhg_fft2_plan_exec on device memory: forward against numpy, then the
    unnormalised inverse returns N times the input; repeats chain."""
    from hhgamd import lib
    rng = np.random.default_rng(7 + shape[1])
    a = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    n = shape[1] * shape[2]
    with lib.FFT2Plan(*shape) as plan, lib.DeviceArray(shape) as x:
        x.upload(a)
        ms = plan.exec(x.address, -1)
        got = x.download()
        ref = np.fft.fft2(a)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert plan.exec(x.address, +1) > 0
        err_back = np.abs(x.download() / n - a).max() / np.abs(a).max()
        # forward twice = N * a(-m) per field (index reversal on both axes)
        x.upload(a)
        plan.exec(x.address, -1, repeats=2)
        rev = np.roll(a[:, ::-1, ::-1], 1, axis=(1, 2)) * n
        err_twice = np.abs(x.download() - rev).max() / np.abs(rev).max()
    print(f"{shape}: forward {err:.1e}, inverse {err_back:.1e}, twice {err_twice:.1e}, {ms:.3f} ms")
    assert err <= 1e-13 and err_back <= 1e-13 and err_twice <= 1e-13


def test_fft2_plan_rejects_bad_memory_before_launching():
    """Host memory, a misaligned address and an allocation smaller than the
    fields are refused on the host; repeats must be positive."""
    from hhgamd import lib
    with lib.FFT2Plan(8, 256, 256) as plan:  # 8 MiB of complex128
        host = np.zeros((8, 256, 256), dtype=np.complex128)
        with pytest.raises(RuntimeError, match="not device memory"):
            plan.exec(host.ctypes.data)
        with lib.DeviceArray((8, 256, 256)) as x:
            with pytest.raises(RuntimeError, match="aligned"):
                plan.exec(x.address + 8)
            with pytest.raises(RuntimeError, match="repeats"):
                plan.exec(x.address, repeats=0)
        with lib.DeviceArray((1 << 16,)) as small:  # 1 MiB
            with pytest.raises(RuntimeError, match="fewer than"):
                plan.exec(small.address)
    with pytest.raises(RuntimeError, match="nch"):
        lib.FFT2Plan(0, 8, 8)


def _run(p, kernel):
    from hhgamd import lib
    with lib.Solve(dict(p, kernel=kernel)) as s:
        s.run()
        k = s.kernel()
        return k, s.current()[1]


@pytest.mark.parametrize("path,nr,pol", [(HMCASE0, (10, 12), (0.0, 1.0, 0.0)), (HALDANE, (12, 16), (0.6, 0.8, 0.0))])
def test_fftform_matches_oracle_fft_form(oracle_lib, path, nr, pol):
    p = make_param(path, nr=nr, tmax=10.0, tmax_field=8.0, pol=pol, intensity=2e12)
    k, J = _run(p, "fftform")
    assert k == 7
    _, Jref = oracle_lib.run(path, p, mode="fft")
    err = rel_l2(J, Jref)
    print(f"FFT form vs oracle FFT form: {err:.2e}")
    assert err <= FP64_TOL


def test_fftform_matches_per_k_kernels(tmp_path):
    """nb <= 4 runs each RK4 stage as three fused launches (forming + rows,
    columns + commutator + inverse columns, inverse rows); nb = 5 takes the
    unfused path (a column of all 2 nb^2 fields no longer fits in LDS)."""
    from hhgamd import tbgen
    paths = {}
    for nb in (3, 4, 5):
        paths[nb] = str(tmp_path / f"syn{nb}_tb.dat")
        tbgen.write_tb(paths[nb], *tbgen.synthetic(nb, seed=5 + nb))
    for pth, nr in [(HALDANE, (24, 30)), (paths[3], (10, 14)), (paths[4], (12, 10)), (paths[5], (8, 10))]:
        p = make_param(pth, nr=nr, tmax=12.0, tmax_field=10.0, pol=(0.8, 0.6, 0.0), phi=(0.3, 0.0, 0.0),
                       intensity=3e12)
        _, Jf = _run(p, "fftform")
        _, Jk = _run(p, "auto")
        err = rel_l2(Jf, Jk)
        print(f"{nr}: FFT form vs per-k {err:.2e}")
        assert err <= FP64_TOL


def test_fftform_at_the_reference_default_grid():
    """400 x 400 (include/settings.h:4): lines over 256 points run one row or
    four columns per tile and the stages take the unfused path; against the
    per-k kernel on a short pulse."""
    p = make_param(HALDANE, nr=(400, 400), tmax=1.2, tmax_field=1.0, pol=(0.6, 0.8, 0.0), intensity=4e12)
    k, Jf = _run(p, "fftform")
    _, Jk = _run(p, "auto")
    err = rel_l2(Jf, Jk)
    print(f"400 x 400: FFT form vs per-k {err:.2e}")
    assert k == 7 and err <= FP64_TOL


def test_fftform_eager_steps_equal_graph_replay(monkeypatch):
    """HHG_FFTFORM_EAGER=1: the step's kernels enqueued directly give bitwise the graph replay's J."""
    p = make_param(HALDANE, nr=(16, 20), tmax=6.0, tmax_field=5.0, pol=(0.6, 0.8, 0.0), intensity=3e12)
    _, Jg = _run(p, "fftform")
    monkeypatch.setenv("HHG_FFTFORM_EAGER", "1")
    _, Je = _run(p, "fftform")
    assert np.array_equal(Je, Jg)


def test_fftform_continues_and_rejects_fp32():
    from hhgamd import lib
    p = make_param(HMCASE0, nr=(8, 8), tmax=4.0, tmax_field=3.0)
    with lib.Solve(dict(p, kernel="fftform")) as s:
        s.run()
        s.run()
        t, J = s.current()
        assert len(t) == s.nt and np.isfinite(J).all()
    with pytest.raises(RuntimeError, match="FP64"):
        with lib.Solve(dict(p, kernel="fftform", precision="fp32")) as s:
            s.run()


def test_fftform_solve_cannot_join_a_batch():
    from hhgamd import lib
    p = make_param(HMCASE0, nr=(8, 8), tmax=2.0, tmax_field=1.5)
    with lib.Solve(dict(p, kernel="fftform")) as s:
        with pytest.raises(RuntimeError, match="runs alone"):
            lib.Batch([s])
