"""numpy restatement of the reference's real-space FFT-convolution time loop.

TEST INFRASTRUCTURE ONLY (tests/ import it as a checker). It is written
independently of ``swe_oracle.cpp`` -- its own parser, numpy.fft for the
transforms, numpy.linalg.eigh for the band basis, and libm ``sinf``/``sqrtf``
through ctypes for the single-precision field -- so that agreement between the
two pins the C++ oracle's FFT, shifts, placement, velocity and stage logic.
Pure numpy, so it is for tiny grids and short runs only.

Follows: src/wannier_tb.cpp:15-113 (parser), src/grid.cpp:64-110 (grids),
src/efield.cpp:30-50 (field), src/velocity.cpp:41-115 (velocity),
src/swe.cpp:49-93,183-233 (setup order, loop, Peierls phase),
src/solver.cpp:86-181,1411-1726 (RK4 and the FFT-convolution stages),
src/fftw_helper.cpp:5-79 (transforms and shifts), src/observable.cpp:16-30 (J).
"""
import ctypes as ct

import numpy as np

_libm = ct.CDLL("libm.so.6")
_libm.sinf.argtypes = [ct.c_float]
_libm.sinf.restype = ct.c_float
_libm.sqrtf.argtypes = [ct.c_float]
_libm.sqrtf.restype = ct.c_float


def parse_tb(path):
    with open(path) as f:
        lines = f.read().split("\n")
    pos = 1
    lat = np.array([[float(x) for x in lines[pos + i].split()] for i in range(3)])
    pos += 3
    nb = int(lines[pos]); nR = int(lines[pos + 1]); pos += 2
    pos += (nR - 1) // 15 + 1
    pos += 1
    R = np.zeros((nR, 3), dtype=int)
    H = np.zeros((nR, nb, nb), dtype=complex)
    r = np.zeros((3, nR, nb, nb), dtype=complex)
    for i in range(nR):
        R[i] = [int(x) for x in lines[pos].split()[:3]]; pos += 1
        for _ in range(nb * nb):
            tok = lines[pos].split(); pos += 1
            m, n = int(tok[0]) - 1, int(tok[1]) - 1
            H[i, n, m] = float(tok[2]) + 1j * float(tok[3])
        pos += 1
    for i in range(nR):
        pos += 1
        for _ in range(nb * nb):
            tok = lines[pos].split(); pos += 1
            m, n = int(tok[0]) - 1, int(tok[1]) - 1
            v = [float(x) for x in tok[2:8]]
            for a in range(3):
                r[a, i, n, m] = v[2 * a] + 1j * v[2 * a + 1]
        pos += 1
    return lat, R, H, r


def fields(prm):
    nt = int(prm["tmax"] / prm["dt"])
    t = np.array([(float(i) * prm["dt"]) * (1.0 / 0.02418884) for i in range(nt)])
    omega = 2.0 * np.pi * 2997.025 / (prm["lambda"] * 10.0)
    omega *= 1.0 / (1.0 / 0.02418884)
    tmf = prm["tmax_field"] * (1.0 / 0.02418884)
    E0 = float(_libm.sqrtf(prm["intensity"] * 1.0 / 3.56e16))
    s = np.asarray(prm["pol_vec"], dtype=float)
    phi = np.asarray(prm["phi_vec"], dtype=float)
    E = np.zeros((3, nt)); A = np.zeros((3, nt))
    for i in range(nt):
        sv = float(np.sin(np.pi * t[i] / tmf))
        env = sv * sv if t[i] < tmf else 0.0  # pow(x, 2) is emitted as x*x by the compiler
        for a in range(3):
            E[a, i] = s[a] * E0 * env * float(_libm.sinf(omega * t[i] + phi[a]))
    dt = t[1] - t[0]
    for i in range(1, nt):
        A[:, i] = A[:, i - 1] - E[:, i - 1] * dt
    return t, E, A


def _fftshift(x):  # shift N//2 on both axes (src/fftw_helper.cpp:36-56)
    return np.roll(x, (x.shape[0] // 2, x.shape[1] // 2), axis=(0, 1))


def _ifftshift(x):  # shift (N+1)//2 (src/fftw_helper.cpp:58-79)
    return np.roll(x, ((x.shape[0] + 1) // 2, (x.shape[1] + 1) // 2), axis=(0, 1))


def run(path, prm):
    """Returns t (nt,) and J (nt, 3) complex, as SWESim_get_current."""
    lat, Rw, Hw, rw = parse_tb(path)
    nb = Hw.shape[1]
    N1, N2 = prm["nr1"], prm["nr2"]
    n1 = np.arange(N1) - N1 // 2
    n2 = np.arange(N2) - N2 // 2
    # supercell fields, channel-major [i, j, i1, i2], SI
    H = np.zeros((nb, nb, N1, N2), dtype=complex)
    X = np.zeros((3, nb, nb, N1, N2), dtype=complex)
    for i1 in range(N1):
        for i2 in range(N2):
            hit = [w for w in range(len(Rw)) if tuple(Rw[w]) == (n1[i1], n2[i2], 0)]
            if hit:
                H[:, :, i1, i2] = Hw[hit[0]]
                X[:, :, :, i1, i2] = -rw[:, hit[0]]
    RA = np.zeros((3, N1, N2))
    for a in range(3):
        RA[a] = n1[:, None] * lat[0, a] + n2[None, :] * lat[1, a]
    Rau = RA * (1.0 / 0.5291772)

    fw = lambda x: np.fft.fft2(_fftshift(x))
    bw = lambda x: _ifftshift(np.fft.ifft2(x))
    # velocity (SI)
    V = np.zeros((3, nb, nb, N1, N2), dtype=complex)
    Hk_ = np.array([[fw(H[i, j]) for j in range(nb)] for i in range(nb)])
    for a in range(3):
        Xk = np.array([[fw(X[a, i, j]) for j in range(nb)] for i in range(nb)])
        comm = np.einsum("ilk,ljk->ijk", Xk.reshape(nb, nb, -1), Hk_.reshape(nb, nb, -1)) \
            - np.einsum("ilk,ljk->ijk", Hk_.reshape(nb, nb, -1), Xk.reshape(nb, nb, -1))
        comm = comm.reshape(nb, nb, N1, N2)
        for i in range(nb):
            for j in range(nb):
                V[a, i, j] = 1j * (RA[a] * H[i, j] - bw(comm[i, j]))
    H = H * 1.0 / 27.2114
    X = X * 1.0 / 0.5291772
    # initial state
    Hk = np.array([[np.fft.fft2(_ifftshift(H[i, j])) for j in range(nb)] for i in range(nb)])
    Hk = Hk.reshape(nb, nb, -1).transpose(2, 0, 1)
    w, U = np.linalg.eigh(Hk)
    HH = np.conj(U.transpose(0, 2, 1)) @ Hk @ U
    occ = (np.real(np.diagonal(HH, axis1=1, axis2=2)) < 0).astype(float)
    rhok = (U * occ[:, None, :]) @ np.conj(U.transpose(0, 2, 1))
    Hback = U @ HH @ np.conj(U.transpose(0, 2, 1))
    to_r = lambda k: _ifftshift(np.fft.ifft2(k.reshape(N1, N2)))
    H = np.array([[to_r(Hback[:, i, j]) for j in range(nb)] for i in range(nb)])
    rho = np.array([[to_r(rhok[:, i, j]) for j in range(nb)] for i in range(nb)])

    t, E, A = fields(prm)
    nt = len(t)
    dt = t[1] - t[0]
    J = np.zeros((nt, 3), dtype=complex)

    def deriv(p, heff, rr):
        hk = np.array([[fw(p * heff[i, j]) for j in range(nb)] for i in range(nb)]).reshape(nb, nb, -1)
        rk = np.array([[fw(rr[i, j]) for j in range(nb)] for i in range(nb)]).reshape(nb, nb, -1)
        c = np.einsum("ilk,ljk->ijk", hk, rk) - np.einsum("ilk,ljk->ijk", rk, hk)
        c = c.reshape(nb, nb, N1, N2)
        return np.array([[-1j * bw(c[i, j]) for j in range(nb)] for i in range(nb)])

    for ti in range(nt):
        p = np.exp(1j * (A[0, ti] * Rau[0] + A[1, ti] * Rau[1] + A[2, ti] * Rau[2]))
        if ti < nt - 1:
            es = [E[:, ti], 0.5 * E[:, ti] + 0.5 * E[:, ti + 1], E[:, ti + 1]]
        else:
            es = [E[:, nt - 1]] * 3
        heff = [H + e[0] * X[0] + e[1] * X[1] + e[2] * X[2] for e in es]
        k1 = deriv(p, heff[0], rho)
        k2 = deriv(p, heff[1], rho + 0.5 * dt * k1)
        k3 = deriv(p, heff[1], rho + 0.5 * dt * k2)
        k4 = deriv(p, heff[2], rho + dt * k3)
        rho = rho + 0.1666666666666 * dt * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
        rt = rho * np.conj(p)
        sl = (slice(None), slice(None), slice(N1 // 2 - 2, N1 // 2 + 2), slice(N2 // 2 - 2, N2 // 2 + 2))
        for a in range(3):
            J[ti, a] = np.sum(V[a][sl] * np.conj(rt[sl]))
    return t, J
