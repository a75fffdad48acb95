"""Writers and generators of Wannier90 ``_tb.dat`` models.

File layout as read by the reference (src/wannier_tb.cpp:15-113): comment,
three lattice rows (Angstrom), nb, nR, degeneracies (15 per line), blank; per
R a line "n1 n2 n3", nb*nb lines "m n Re Im" (m fastest) in eV and a blank;
then per R the position blocks "m n Re(x) Im(x) Re(y) Im(y) Re(z) Im(z)" in
Angstrom. A line "m n" holds <m,0|O|n,R>.

Blocks passed to :func:`write_tb` use the file convention: ``H[iR, m, n]`` is
<m,0|H|n,R>.
"""
import numpy as np

HMCASE0_LATTICE = np.array([[2.130422493, -1.23, 0.0],
                            [2.130422493, 1.23, 0.0],
                            [0.0, 0.0, 15.0]])
B_SITE_X = 1.42028  # Angstrom, orbital 2 of the honeycomb models


def write_tb(path, lattice, R, H, r, comment="written by hhgamd.tbgen"):
    """Write a model. R: (nR, 3) ints; H: (nR, nb, nb) eV; r: (3, nR, nb, nb) Angstrom."""
    R = np.asarray(R, dtype=int)
    H = np.asarray(H, dtype=complex)
    r = np.asarray(r, dtype=complex)
    nR, nb = H.shape[0], H.shape[1]
    out = [comment]
    for a in lattice:
        out.append(" ".join(f"{x:.18e}" for x in a))
    out.append(str(nb))
    out.append(str(nR))
    deg = ["1"] * nR
    for i in range(0, nR, 15):
        out.append("  ".join(deg[i:i + 15]) + "   ")
    out.append("")
    for i in range(nR):
        out.append(f"{R[i, 0]:5d}{R[i, 1]:5d}{R[i, 2]:5d}")
        for n in range(nb):
            for m in range(nb):
                v = H[i, m, n]
                out.append(f"{m + 1:3d}{n + 1:4d}  {v.real:.17e}  {v.imag:.17e}")
        out.append("")
    for i in range(nR):
        out.append(f"{R[i, 0]:5d}{R[i, 1]:5d}{R[i, 2]:5d}")
        for n in range(nb):
            for m in range(nb):
                vals = " ".join(f"{r[a, i, m, n].real:.17e} {r[a, i, m, n].imag:.17e}" for a in range(3))
                out.append(f"{m + 1:5d}{n + 1:5d}  {vals}")
        out.append("")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def shell(nmax):
    """R vectors n1, n2 in [-nmax, nmax], n3 = 0, row-major."""
    return np.array([[a, b, 0] for a in range(-nmax, nmax + 1) for b in range(-nmax, nmax + 1)], dtype=int)


def haldane(mass=0.3, t1=2.8, t2=0.1, phi=np.pi / 2, nmax=2):
    """Haldane-like honeycomb on the hmcase0 lattice.

    Orbital 1 (A) at the origin, orbital 2 (B) at x = 1.42028 Angstrom; on-site
    +mass (A) and -mass (B); nearest-neighbour hopping -t1 on the three A-B
    bonds (hmcase0's sign convention); complex next-nearest-neighbour hopping
    t2 e^{+i phi nu} on A and t2 e^{-i phi nu} on B, with nu = +1 along a1,
    a2-a1, -a2 and -1 along the opposite vectors. Zero blocks fill the shell
    |n_i| <= nmax. Returns (lattice, R, H, r) in the file convention.
    """
    R = shell(nmax)
    nR = len(R)
    H = np.zeros((nR, 2, 2), dtype=complex)
    r = np.zeros((3, nR, 2, 2), dtype=complex)
    idx = {tuple(x[:2]): i for i, x in enumerate(R)}
    H[idx[(0, 0)], 0, 0] = mass
    H[idx[(0, 0)], 1, 1] = -mass
    for n in [(0, 0), (-1, 0), (0, -1)]:          # <A,0|H|B,R>: B at tau_B + R next to A
        H[idx[n], 0, 1] = -t1
        H[idx[(-n[0], -n[1])], 1, 0] = -t1
    for n, nu in [((1, 0), 1), ((-1, 1), 1), ((0, -1), 1), ((-1, 0), -1), ((1, -1), -1), ((0, 1), -1)]:
        H[idx[n], 0, 0] = t2 * np.exp(1j * phi * nu)
        H[idx[n], 1, 1] = t2 * np.exp(-1j * phi * nu)
    r[0, idx[(0, 0)], 1, 1] = B_SITE_X
    return HMCASE0_LATTICE.copy(), R, H, r


def synthetic(nb, seed, nmax=2, amp=2.0, decay=1.5, pos_amp=0.5, gap=3.0, lattice=None):
    """Seeded random Hermitian Wannier model over the shell |n_i| <= nmax.

    |H_R| ~ amp * exp(-|R| / decay) eV with H_{-R} = H_R^dagger; the first nb/2
    orbitals get +gap/2 and the others -gap/2 on site so that E_F = 0 sits in a
    gap for moderate amp. Position blocks ~ pos_amp * exp(-|R| / decay)
    Angstrom with r_{-R} = r_R^dagger, plus real orbital centres at R = 0.
    Generator: numpy PCG64 with the given seed.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    lat = HMCASE0_LATTICE.copy() if lattice is None else np.asarray(lattice, dtype=float)
    R = shell(nmax)
    nR = len(R)
    idx = {tuple(x[:2]): i for i, x in enumerate(R)}
    H = np.zeros((nR, nb, nb), dtype=complex)
    r = np.zeros((3, nR, nb, nb), dtype=complex)
    done = set()
    for i, n in enumerate(R):
        key = (n[0], n[1])
        if key in done:
            continue
        neg = idx[(-n[0], -n[1])]
        dist = np.linalg.norm(n[0] * lat[0] + n[1] * lat[1])
        scale = amp * np.exp(-dist / decay)
        pscale = pos_amp * np.exp(-dist / decay)
        M = scale * (rng.standard_normal((nb, nb)) + 1j * rng.standard_normal((nb, nb))) / np.sqrt(2 * nb)
        P = pscale * (rng.standard_normal((3, nb, nb)) + 1j * rng.standard_normal((3, nb, nb))) / np.sqrt(2)
        if neg == i:
            M = 0.5 * (M + M.conj().T)
            P = 0.5 * (P + P.conj().transpose(0, 2, 1))
            M = M + np.diag([gap / 2 if o < nb // 2 else -gap / 2 for o in range(nb)])
            centres = rng.uniform(0.0, 1.0, (nb, 2)) @ lat[:2, :2]
            for o in range(nb):
                P[0, o, o] = centres[o, 0]
                P[1, o, o] = centres[o, 1]
                P[2, o, o] = 0.0
            H[i] = M
            r[:, i] = P
        else:
            H[i] = M
            H[neg] = M.conj().T
            r[:, i] = P
            r[:, neg] = P.conj().transpose(0, 2, 1)
        done.add(key)
        done.add((-n[0], -n[1]))
    return lat, R, H, r
