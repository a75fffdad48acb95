// sbe_dumps.cpp -- the text dumps of SWESim::test_files (src/swe.cpp:131-181).
//
// Files, formats and units follow the reference:
//   test/Rvecs_{SI,AU}.dat, test/Kvecs_{SI,AU}.dat    unit tag line, then "x y z"
//       per supercell point (src/grid.cpp:183-214); R = sum_d n_d a_d with
//       n_d = i_d - nr_d/2, k = sum_d (m_d / nr_d) b_d with m_d = i_d
//       (src/grid.cpp:64-107), AU: R x (1/0.5291772), k / (1/0.5291772)
//   MatrixField dumps (src/matrix_field.cpp:28-53): unit tag line, then
//       "i |z| arg(z)" per point i and element (row-major block element)
//       test/hamiltonian_{SI,AU}.dat        H(R), eV / Hartree
//       test/hamiltonian_k_AU.dat           H_k = FFT(ifftshift(H)) (src/operator.cpp:41-64)
//       test/hamiltonian_r_AU.dat           back to R: ifftshift(IFFT(H_k)) / N
//       test/Uk_AU.dat                      eigenvectors per k, U[i][j] = component i of
//                                           eigenvector j, ascending eigenvalues
//       test/hamiltonian_k_H_AU.dat         U^dagger H_k U
//       test/hamiltonian_k_W_AU.dat         U (U^dagger H_k U) U^dagger
//       test/rho_k_H_AU.dat                 diag(occupation), occupation = [(U^dagger H_k U)_nn < 0]
//       test/rho_k_W_AU.dat                 U diag(occupation) U^dagger
//       test/rho_r_W_AU.dat                 ifftshift(IFFT(rho_k_W)) / N
//       test/{x,y,z}_bc.dat                 rbc = -r in bohr on the supercell
//   fields/efield.dat, fields/afield.dat  unit-system integer (1 = AU), then
//       "t Ex Ey Ez" / "t Ax Ay Az" per step (src/efield.cpp:91-113)
// Eigenvectors carry an arbitrary phase per column (the reference's zheevd
// fixes another one), so Uk_AU.dat matches the reference only up to column
// phases; every other file is independent of that choice. Unlike the
// reference, writing the dumps leaves the solve's state unchanged.
#include <algorithm>
#include <cmath>
#include <complex>
#include <filesystem>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "sbe_host.h"

namespace hhg {

namespace {

struct Field {  // nb*nb elements per supercell point, natural order (n = i - N/2)
    int N, nb2;
    std::vector<cplx> v;
    Field(int N_, int nb2_) : N(N_), nb2(nb2_), v((size_t)N_ * nb2_, 0.0) {}
    cplx& at(int i, int e) { return v[(size_t)i * nb2 + e]; }
};

void write_matrix(const std::string& path, const char* tag, Field& f) {
    std::ofstream out(path);
    if (!out) throw std::runtime_error("cannot write " + path);
    out << tag << std::endl;
    for (int i = 0; i < f.N; ++i)
        for (int e = 0; e < f.nb2; ++e) {
            const cplx z = f.at(i, e);
            out << i << " " << std::abs(z) << " " << std::atan2(z.imag(), z.real()) << std::endl;
        }
}

void write_vectors(const std::string& path, const char* tag, const std::vector<double>& xyz) {
    std::ofstream out(path);
    if (!out) throw std::runtime_error("cannot write " + path);
    out << tag << std::endl;
    for (size_t i = 0; i < xyz.size() / 3; ++i) out << xyz[3 * i] << " " << xyz[3 * i + 1] << " " << xyz[3 * i + 2] << std::endl;
}

// 2-D transform of every element between the natural R order and the FFT
// order of k: forward = FFT(ifftshift(f)), backward = ifftshift(IFFT(f)) / N
// (even N: the shift is a half-period roll on both axes).
Field transform(const Field& f, int N1, int N2, bool forward) {
    const int N = N1 * N2, nb2 = f.nb2;
    std::vector<double> buf(2 * (size_t)nb2 * N);
    auto roll = [&](int i) { int i1 = i / N2, i2 = i % N2; return ((i1 + N1 / 2) % N1) * N2 + (i2 + N2 / 2) % N2; };
    for (int e = 0; e < nb2; ++e)
        for (int i = 0; i < N; ++i) {
            const cplx z = f.v[(size_t)(forward ? roll(i) : i) * nb2 + e];
            buf[2 * ((size_t)e * N + i)] = z.real();
            buf[2 * ((size_t)e * N + i) + 1] = z.imag();
        }
    fft2_host(buf.data(), nb2, N1, N2, forward ? -1 : +1);
    Field out(N, nb2);
    for (int e = 0; e < nb2; ++e)
        for (int i = 0; i < N; ++i) {
            cplx z(buf[2 * ((size_t)e * N + i)], buf[2 * ((size_t)e * N + i) + 1]);
            if (!forward) z = cplx(z.real() / N, z.imag() / N);
            out.v[(size_t)(forward ? i : roll(i)) * nb2 + e] = z;
        }
    return out;
}

}  // namespace

void write_test_files(const TBModel& m, const HostSolve& hs, const std::string& test_dir,
                      const std::string& fields_dir) {
    const SolveDesc& d = hs.d;
    const int nb = d.nb, nb2 = nb * nb, N1 = d.N1, N2 = d.N2, N = d.N;
    std::filesystem::create_directories(test_dir);
    std::filesystem::create_directories(fields_dir);
    const std::string T = test_dir + "/";

    // grid (src/grid.cpp:64-107), reciprocal vectors b_i = 2 pi (a_j x a_k) / V
    const double(&a)[3][3] = m.lat;
    auto cross = [](const double* u, const double* v, double* w) {
        w[0] = u[1] * v[2] - u[2] * v[1]; w[1] = u[2] * v[0] - u[0] * v[2]; w[2] = u[0] * v[1] - u[1] * v[0];
    };
    double b[3][3], c[3];
    cross(a[1], a[2], c);
    const double vol = a[0][0] * c[0] + a[0][1] * c[1] + a[0][2] * c[2];
    for (int i = 0; i < 3; ++i) {
        cross(a[(i + 1) % 3], a[(i + 2) % 3], c);
        for (int x = 0; x < 3; ++x) b[i][x] = 2.0 * M_PI * c[x] / vol;
    }
    std::vector<double> R(3 * (size_t)N), K(3 * (size_t)N);
    for (int i1 = 0; i1 < N1; ++i1)
        for (int i2 = 0; i2 < N2; ++i2) {
            const int i = i1 * N2 + i2, n1 = -N1 / 2 + i1, n2 = -N2 / 2 + i2;
            for (int x = 0; x < 3; ++x) {
                R[3 * (size_t)i + x] = n1 * a[0][x] + n2 * a[1][x] + 0 * a[2][x];
                K[3 * (size_t)i + x] = (double)i1 / (double)N1 * b[0][x] + (double)i2 / (double)N2 * b[1][x] +
                                       0.0 / 1.0 * b[2][x];
            }
        }
    write_vectors(T + "Rvecs_SI.dat", "_SI_", R);
    write_vectors(T + "Kvecs_SI.dat", "_SI_", K);
    for (size_t i = 0; i < R.size(); ++i) { R[i] *= 1.0 / 0.5291772; K[i] /= 1.0 / 0.5291772; }
    write_vectors(T + "Rvecs_AU.dat", "_AU_", R);
    write_vectors(T + "Kvecs_AU.dat", "_AU_", K);

    // operators placed on the supercell at the integer R of every Wannier block
    // (points outside the supercell are dropped, as by the placement of the solve)
    Field Hsi(N, nb2), H(N, nb2), X[3] = {Field(N, nb2), Field(N, nb2), Field(N, nb2)};
    for (int w = 0; w < m.nR; ++w) {
        const int n1 = m.R[3 * w], n2 = m.R[3 * w + 1], n3 = m.R[3 * w + 2];
        if (n3 != 0 || n1 < -N1 / 2 || n1 > N1 / 2 - 1 || n2 < -N2 / 2 || n2 > N2 / 2 - 1) continue;
        const int i = (n1 + N1 / 2) * N2 + (n2 + N2 / 2);
        for (int e = 0; e < nb2; ++e) {
            Hsi.at(i, e) = m.H[(size_t)w * nb2 + e];
            H.at(i, e) = m.H[(size_t)w * nb2 + e] * (1.0 / 27.2114);
            for (int x = 0; x < 3; ++x) X[x].at(i, e) = (-m.r[x][(size_t)w * nb2 + e]) * (1.0 / 0.5291772);
        }
    }
    write_matrix(T + "hamiltonian_SI.dat", "_SI_", Hsi);
    write_matrix(T + "hamiltonian_AU.dat", "_AU_", H);
    Field Hk = transform(H, N1, N2, true);
    write_matrix(T + "hamiltonian_k_AU.dat", "_AU_", Hk);
    Field Hr = transform(Hk, N1, N2, false);
    write_matrix(T + "hamiltonian_r_AU.dat", "_AU_", Hr);

    // band basis, the H-gauge Hamiltonian and the T = 0 state per k
    Field U(N, nb2), HH(N, nb2), HW(N, nb2), RH(N, nb2), RW(N, nb2);
    std::vector<cplx> A(nb2), Q, UH(nb2);
    for (int k = 0; k < N; ++k) {
        for (int e = 0; e < nb2; ++e) A[e] = Hk.at(k, e);
        std::vector<cplx> a0 = A;
        herm_jacobi(nb, a0, Q);
        std::vector<int> order(nb);
        for (int j = 0; j < nb; ++j) order[j] = j;
        std::sort(order.begin(), order.end(), [&](int x, int y) { return a0[(size_t)x * nb + x].real() < a0[(size_t)y * nb + y].real(); });
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j) U.at(k, i * nb + j) = Q[(size_t)i * nb + order[j]];
        auto u = [&](int i, int j) { return U.at(k, i * nb + j); };
        for (int i = 0; i < nb; ++i)  // UH = U^dagger H
            for (int j = 0; j < nb; ++j) {
                cplx s = 0;
                for (int l = 0; l < nb; ++l) s += std::conj(u(l, i)) * A[(size_t)l * nb + j];
                UH[(size_t)i * nb + j] = s;
            }
        for (int i = 0; i < nb; ++i)  // HH = U^dagger H U
            for (int j = 0; j < nb; ++j) {
                cplx s = 0;
                for (int l = 0; l < nb; ++l) s += UH[(size_t)i * nb + l] * u(l, j);
                HH.at(k, i * nb + j) = s;
            }
        for (int i = 0; i < nb; ++i)  // back to the W gauge and the T = 0 state
            for (int j = 0; j < nb; ++j) {
                cplx hw = 0, rw = 0;
                for (int p = 0; p < nb; ++p)
                    for (int q = 0; q < nb; ++q) hw += u(i, p) * HH.at(k, p * nb + q) * std::conj(u(j, q));
                for (int n = 0; n < nb; ++n)
                    if (HH.at(k, n * nb + n).real() < 0.0) rw += u(i, n) * std::conj(u(j, n));
                HW.at(k, i * nb + j) = hw;
                RW.at(k, i * nb + j) = rw;
                RH.at(k, i * nb + j) = (i == j && HH.at(k, i * nb + i).real() < 0.0) ? 1.0 : 0.0;
            }
    }
    write_matrix(T + "Uk_AU.dat", "_AU_", U);
    write_matrix(T + "hamiltonian_k_H_AU.dat", "_AU_", HH);
    write_matrix(T + "hamiltonian_k_W_AU.dat", "_AU_", HW);
    write_matrix(T + "rho_k_H_AU.dat", "_AU_", RH);
    write_matrix(T + "rho_k_W_AU.dat", "_AU_", RW);
    Field Rr = transform(RW, N1, N2, false);
    write_matrix(T + "rho_r_W_AU.dat", "_AU_", Rr);

    // fields in au (src/efield.cpp:91-113)
    {
        std::ofstream fe(fields_dir + "/efield.dat"), fa(fields_dir + "/afield.dat");
        if (!fe || !fa) throw std::runtime_error("cannot write the field files in " + fields_dir);
        fe << 1 << std::endl;
        fa << 1 << std::endl;
        for (int i = 0; i < d.nt; ++i) {
            fe << hs.t[i] << " " << hs.E[0][i] << " " << hs.E[1][i] << " " << hs.E[2][i] << std::endl;
            fa << hs.t[i] << " " << hs.A[0][i] << " " << hs.A[1][i] << " " << hs.A[2][i] << std::endl;
        }
    }
    write_matrix(T + "x_bc.dat", "_AU_", X[0]);
    write_matrix(T + "y_bc.dat", "_AU_", X[1]);
    write_matrix(T + "z_bc.dat", "_AU_", X[2]);
}

}  // namespace hhg
