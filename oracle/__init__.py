"""CPU FP64 oracle for the HHG-Solids k-space time loop (ctypes wrapper).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` -- as the checker, never as the thing
measured for the GPU path and never by the product package.

Parameters use the same dictionary keys as the reference's ctypes driver
(``pySWE/pyswe.py:59-98``): nr1, nr2, nr3, tmax, dt, intensity, lambda,
tmax_field, pol_vec, phi_vec. Times are in the reference API's units (fs,
converted with 1/0.02418884 inside, ``src/grid.cpp:133-135``).
"""
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C")
_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C")


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ct.CDLL(LIB_PATH)
        L.oracle_last_error.restype = ct.c_char_p
        L.oracle_nt.argtypes = [_I, _D, ct.POINTER(ct.c_int)]
        L.oracle_model_dims.argtypes = [ct.c_char_p, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]
        L.oracle_model.argtypes = [ct.c_char_p, _D, _I, _D, _D]
        L.oracle_fields.argtypes = [_I, _D, _D, _D, _D]
        L.oracle_hk.argtypes = [ct.c_char_p, _I, _D, _D]
        L.oracle_rho0.argtypes = [ct.c_char_p, _I, _D, ct.c_int, _D]
        L.oracle_velocity_window.argtypes = [ct.c_char_p, _I, _D, ct.c_int, _D]
        L.oracle_run.argtypes = [ct.c_char_p, _I, _D, ct.c_int, ct.c_int, _D, _D]
        L.oracle_run_from.argtypes = [ct.c_char_p, _I, _D, _D, ct.c_int, _D, _D]
        L.oracle_propagate_k.argtypes = [ct.c_char_p, _I, _D, ct.c_int, _I, ct.c_int, _D]
        L.oracle_fft2.argtypes = [_D, ct.c_int, ct.c_int, ct.c_int]
        L.oracle_eigh.argtypes = [ct.c_int, _D, _D, _D]
        L.oracle_equilibrium_rho.argtypes = [ct.c_int, _D, _D]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().oracle_last_error().decode())


def pack(prm):
    nr = np.array([prm["nr1"], prm["nr2"], prm.get("nr3", 1)], dtype=np.int32)
    dp = np.array([prm["tmax"], prm["dt"], prm["intensity"], prm["lambda"], prm["tmax_field"],
                   *np.asarray(prm["pol_vec"], dtype=np.float64),
                   *np.asarray(prm["phi_vec"], dtype=np.float64)], dtype=np.float64)
    return nr, dp


def nt(prm):
    nr, dp = pack(prm)
    out = ct.c_int()
    _check(lib().oracle_nt(nr, dp, ct.byref(out)))
    return out.value


def model_dims(path):
    nb, nR = ct.c_int(), ct.c_int()
    _check(lib().oracle_model_dims(path.encode(), ct.byref(nb), ct.byref(nR)))
    return nb.value, nR.value


def model(path):
    """Parsed blocks (SI): lat (3,3), R (nR,3), H (nR,nb,nb), r (3,nR,nb,nb); H[iR][n][m] = file line 'm n'."""
    nb, nR = model_dims(path)
    lat = np.zeros(9)
    R = np.zeros(3 * nR, dtype=np.int32)
    H = np.zeros(2 * nR * nb * nb)
    r = np.zeros(6 * nR * nb * nb)
    _check(lib().oracle_model(path.encode(), lat, R, H, r))
    return (lat.reshape(3, 3), R.reshape(nR, 3), H.view(np.complex128).reshape(nR, nb, nb),
            r.view(np.complex128).reshape(3, nR, nb, nb))


def fields(prm):
    n = nt(prm)
    nr, dp = pack(prm)
    t = np.zeros(n); E = np.zeros(3 * n); A = np.zeros(3 * n)
    _check(lib().oracle_fields(nr, dp, t, E, A))
    return t, E.reshape(3, n), A.reshape(3, n)


def hk(path, prm):
    nb, _ = model_dims(path)
    nr, dp = pack(prm)
    N = int(nr[0]) * int(nr[1])
    out = np.zeros(2 * N * nb * nb)
    _check(lib().oracle_hk(path.encode(), nr, dp, out))
    return out.view(np.complex128).reshape(N, nb, nb)


def rho0(path, prm, mode="perk"):
    nb, _ = model_dims(path)
    nr, dp = pack(prm)
    N = int(nr[0]) * int(nr[1])
    out = np.zeros(2 * N * nb * nb)
    _check(lib().oracle_rho0(path.encode(), nr, dp, 0 if mode == "perk" else 1, out))
    return out.view(np.complex128).reshape(N, nb, nb)


def velocity_window(path, prm, mode="direct"):
    """v_a(n) (eV*Angstrom) at n1, n2 in {-2..1}: shape (3, 4, 4, nb, nb)."""
    nb, _ = model_dims(path)
    nr, dp = pack(prm)
    out = np.zeros(2 * 3 * 16 * nb * nb)
    _check(lib().oracle_velocity_window(path.encode(), nr, dp, 0 if mode == "direct" else 1, out))
    return out.view(np.complex128).reshape(3, 4, 4, nb, nb)


def run(path, prm, mode="perk", nthreads=None):
    """Full solve. Returns t (nt,) in au and J (nt, 3) complex128 (eV*Angstrom units)."""
    n = nt(prm)
    nr, dp = pack(prm)
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    t = np.zeros(n)
    J = np.zeros(6 * n)
    _check(lib().oracle_run(path.encode(), nr, dp, 0 if mode == "perk" else 1, int(nthreads), t, J))
    return t, J.view(np.complex128).reshape(n, 3)


def run_from(path, prm, rho_init, nthreads=None):
    """Per-k solve started from the states rho_init (N, nb, nb) instead of the
    initial state (a continued run). Returns t (nt,) and J (nt, 3)."""
    n = nt(prm)
    nr, dp = pack(prm)
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1)
    rho = np.ascontiguousarray(rho_init, dtype=np.complex128)
    if rho.shape[0] != int(nr[0]) * int(nr[1]):
        raise ValueError("rho_init must hold one state per k-point")
    t = np.zeros(n)
    J = np.zeros(6 * n)
    _check(lib().oracle_run_from(path.encode(), nr, dp, rho.view(np.float64).reshape(-1), int(nthreads), t, J))
    return t, J.view(np.complex128).reshape(n, 3)


def propagate_k(path, prm, kidx, nthreads=None):
    """Final rho_k (after all nt RK4 steps, per-k form) of the listed k-points
    k = m1*nr2 + m2: shape (len(kidx), nb, nb) complex128."""
    nb, _ = model_dims(path)
    nr, dp = pack(prm)
    kidx = np.ascontiguousarray(kidx, dtype=np.int32)
    if nthreads is None:
        nthreads = min(16, os.cpu_count() or 1, max(1, len(kidx)))
    out = np.zeros(2 * len(kidx) * nb * nb)
    _check(lib().oracle_propagate_k(path.encode(), nr, dp, len(kidx), kidx, int(nthreads), out))
    return out.view(np.complex128).reshape(len(kidx), nb, nb)


def fft2(a, sign=-1):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    n1, n2 = a.shape
    buf = a.view(np.float64).reshape(-1)
    _check(lib().oracle_fft2(buf, n1, n2, sign))
    return a


def eigh(A):
    A = np.ascontiguousarray(A, dtype=np.complex128)
    n = A.shape[0]
    w = np.zeros(n)
    U = np.zeros((n, n), dtype=np.complex128)
    _check(lib().oracle_eigh(n, A.view(np.float64).reshape(-1), w, U.view(np.float64).reshape(-1)))
    return w, U


def equilibrium_rho(Hk):
    Hk = np.ascontiguousarray(Hk, dtype=np.complex128)
    nb = Hk.shape[0]
    rho = np.zeros((nb, nb), dtype=np.complex128)
    _check(lib().oracle_equilibrium_rho(nb, Hk.view(np.float64).reshape(-1), rho.view(np.float64).reshape(-1)))
    return rho
