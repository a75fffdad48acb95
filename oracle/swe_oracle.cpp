// swe_oracle.cpp -- CPU FP64 oracle for the HHG-Solids k-space time loop.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker, never the product: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load the
// library built from it (oracle/liboracle.so). The product library
// (libwannier.so) neither links nor calls it.
//
// It restates the reference's semantics twice, independently:
//
//   fftform  The reference's own algorithm: rho held in real space on the
//            nr1 x nr2 supercell; every RK4 stage forward-FFTs p(R)*Heff(R)
//            and rho'(R) per orbital channel, forms [H_k, rho_k] per k and
//            inverse-FFTs back (src/solver.cpp:1428-1726), with the same
//            fftshift/ifftshift conventions (src/fftw_helper.cpp:36-79) and
//            the 4x4 real-space current window (src/observable.cpp:16-30).
//            Single-threaded, like the reference. The FFT is written here
//            (mixed radix), not FFTW.
//   perk     The algebraically equal per-k form: every k is an independent
//            nb x nb RK4 with H^(k) = sum_R [H_R + E.rbc_R] e^{-i(k-A).R} and
//            J = (1/N) sum_k sum_ij V^W_ij(k-A) conj(rho_k,ij), where V^W is
//            the Fourier sum of v(R) over the same 4x4 window.
//
// Parity pins (tests/): fftform vs perk (<=1e-13), fftform vs a numpy
// restatement using numpy.fft, FFT vs numpy.fft, Jacobi eigenpairs vs
// numpy.linalg.eigh, fields vs libm sinf/sqrtf called through ctypes,
// honeycomb closed-form bands, projector invariants, RK4 Taylor closed form.
//
// Deliberate difference from the reference (both forms): Wannier blocks are
// placed at exact integer R. The reference solves fractional coordinates by
// Cramer's rule and truncates with (int) (src/hamiltonian.cpp:22-26,
// src/berry_connection.cpp:33-37), which misplaces |n|=2 shells; for models
// whose R all have |n_i| <= 1 both placements agree.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math).

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>
#include <stdexcept>

typedef std::complex<double> cd;

static thread_local std::string g_err;

// ---------------------------------------------------------------- model ----
// Follows src/wannier_tb.cpp:15-113: line-based W90 _tb.dat parser. A line
// "m n re im" is stored at [n-1][m-1] (the file's first orbital index is the
// fast one, src/wannier_tb.cpp:73-75,101-102).
struct TB {
    double lat[3][3];
    int nb = 0, nR = 0;
    std::vector<int> R;          // nR*3
    std::vector<cd> H;           // nR*nb*nb, eV
    std::vector<cd> r[3];        // nR*nb*nb, Angstrom (x, y, z position blocks)
};

static TB parse_tb(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open tb file: " + path);
    TB tb;
    std::string line;
    std::getline(f, line);  // comment
    for (int i = 0; i < 3; ++i) {
        std::getline(f, line);
        std::istringstream is(line);
        is >> tb.lat[i][0] >> tb.lat[i][1] >> tb.lat[i][2];
    }
    std::getline(f, line); tb.nb = std::stoi(line);
    std::getline(f, line); tb.nR = std::stoi(line);
    if (tb.nb <= 0 || tb.nR <= 0) throw std::runtime_error("bad nb/nR");
    int ndeg = (tb.nR - 1) / 15 + 1;
    for (int i = 0; i < ndeg; ++i) std::getline(f, line);  // degeneracies: read and ignored
    std::getline(f, line);                                   // blank
    const int nb = tb.nb, nb2 = nb * nb;
    tb.R.assign(3 * tb.nR, 0);
    tb.H.assign((size_t)tb.nR * nb2, cd(0, 0));
    for (int a = 0; a < 3; ++a) tb.r[a].assign((size_t)tb.nR * nb2, cd(0, 0));
    for (int i = 0; i < tb.nR; ++i) {
        std::getline(f, line);
        std::istringstream is(line);
        is >> tb.R[3 * i] >> tb.R[3 * i + 1] >> tb.R[3 * i + 2];
        for (int j = 0; j < nb2; ++j) {
            std::getline(f, line);
            std::istringstream js(line);
            int m, n; double re, im;
            if (!(js >> m >> n >> re >> im)) throw std::runtime_error("bad hopping line");
            if (m < 1 || m > nb || n < 1 || n > nb) throw std::runtime_error("orbital index out of range");
            tb.H[(size_t)i * nb2 + (n - 1) * nb + (m - 1)] = cd(re, im);
        }
        std::getline(f, line);
    }
    for (int i = 0; i < tb.nR; ++i) {
        std::getline(f, line);  // R line of the position block: not parsed (src/wannier_tb.cpp:87-88)
        for (int j = 0; j < nb2; ++j) {
            std::getline(f, line);
            std::istringstream js(line);
            int m, n; double v[6];
            if (!(js >> m >> n >> v[0] >> v[1] >> v[2] >> v[3] >> v[4] >> v[5]))
                throw std::runtime_error("bad position line");
            if (m < 1 || m > nb || n < 1 || n > nb) throw std::runtime_error("orbital index out of range");
            size_t idx = (size_t)i * nb2 + (n - 1) * nb + (m - 1);
            for (int a = 0; a < 3; ++a) tb.r[a][idx] = cd(v[2 * a], v[2 * a + 1]);
        }
        std::getline(f, line);
    }
    return tb;
}

// ------------------------------------------------------------ parameters ----
struct Params {
    int nr1, nr2, nr3;
    double tmax, dt, intensity, lambda_nm, tmax_field, s[3], phi[3];
    int nt;
};

static Params make_params(const int* nr, const double* dp) {
    Params p;
    p.nr1 = nr[0]; p.nr2 = nr[1]; p.nr3 = nr[2];
    p.tmax = dp[0]; p.dt = dp[1]; p.intensity = dp[2]; p.lambda_nm = dp[3]; p.tmax_field = dp[4];
    for (int a = 0; a < 3; ++a) { p.s[a] = dp[5 + a]; p.phi[a] = dp[8 + a]; }
    p.nt = (int)(p.tmax / p.dt);  // src/swe.cpp:52
    if (p.nr3 != 1) throw std::runtime_error("nr3 must be 1 (2-D FFT only)");
    if (p.nr1 < 4 || p.nr2 < 4 || (p.nr1 & 1) || (p.nr2 & 1))
        throw std::runtime_error("nr1, nr2 must be even and >= 4");
    if (p.nt < 2) throw std::runtime_error("nt = tmax/dt must be >= 2");
    return p;
}

// ----------------------------------------------------------------- fields ----
// src/efield.cpp:7-50 with the unit handling of src/efield.cpp:66-77 and
// src/grid.cpp:64-67,133-135. The macros of include/constants.h expand
// without parentheses; the expressions below keep their evaluation order.
// E0 and sin() are single precision (sqrtf, sinf) by construction.
struct Fields { std::vector<double> t, E[3], A[3]; double dt; };

static Fields make_fields(const Params& p) {
    Fields F;
    const int nt = p.nt;
    F.t.resize(nt);
    for (int i = 0; i < nt; ++i) { F.t[i] = (double)i * p.dt; F.t[i] *= 1.0 / 0.02418884; }
    double lam = p.lambda_nm * 10.0;
    double omega = 2.0 * M_PI * 2997.025 / lam;
    double tmf = p.tmax_field;
    tmf *= 1.0 / 0.02418884;
    omega *= 1.0 / (1.0 / 0.02418884);
    double E0 = sqrtf(p.intensity * 1.0 / 3.56e16);
    double dt = F.t[1] - F.t[0];
    F.dt = dt;
    for (int a = 0; a < 3; ++a) { F.E[a].assign(nt, 0.0); F.A[a].assign(nt, 0.0); }
    for (int i = 0; i < nt; ++i) {
        double ti = F.t[i];
        double env = 0.0;
        if (ti < tmf) { double sv = sin(M_PI * ti / tmf); env = sv * sv; }
        for (int a = 0; a < 3; ++a)
            F.E[a][i] = p.s[a] * E0 * env * sinf(omega * ti + p.phi[a]);
    }
    for (int i = 1; i < nt; ++i)
        for (int a = 0; a < 3; ++a) F.A[a][i] = F.A[a][i - 1] - F.E[a][i - 1] * dt;
    return F;
}

// Stage fields of one step (src/solver.cpp:109-148): t_i, the average of t_i
// and t_{i+1}, and t_{i+1}; the last step uses E[nt-1] for all three.
static void stage_fields(const Fields& F, int ti, int nt, double e[3][3]) {
    for (int a = 0; a < 3; ++a) {
        if (ti < nt - 1) {
            e[0][a] = F.E[a][ti];
            e[1][a] = 0.5 * F.E[a][ti] + 0.5 * F.E[a][ti + 1];
            e[2][a] = F.E[a][ti + 1];
        } else {
            e[0][a] = e[1][a] = e[2][a] = F.E[a][nt - 1];
        }
    }
}

// ------------------------------------------------------------------- FFT ----
// Mixed-radix decimation-in-time FFT, unnormalised, sign -1 = forward
// (FFTW_FORWARD), +1 = backward. Twiddles from long-double cos/sin.
struct FFT1 {
    int n = 0, sign = -1;
    std::vector<int> fac;
    std::vector<cd> tw;  // tw[k] = exp(sign*2*pi*i*k/n)
    void init(int n_, int sign_) {
        n = n_; sign = sign_; fac.clear();
        int m = n;
        for (int p : {4, 2, 3, 5}) while (m % p == 0) { fac.push_back(p); m /= p; }
        for (int p = 7; m > 1; p += 2) while (m % p == 0) { fac.push_back(p); m /= p; }
        tw.resize(n);
        for (int k = 0; k < n; ++k) {
            long double ang = (long double)sign * 2.0L * 3.141592653589793238462643383279502884L * k / n;
            tw[k] = cd((double)cosl(ang), (double)sinl(ang));
        }
    }
    void rec(const cd* in, int is, cd* out, int m, int fi, int tws) const {
        if (m == 1) { out[0] = in[0]; return; }
        int p = fac[fi], q = m / p;
        for (int r = 0; r < p; ++r) rec(in + (size_t)r * is, is * p, out + (size_t)r * q, q, fi + 1, tws * p);
        cd ybuf[64];
        std::vector<cd> ydyn;
        cd* y = ybuf;
        if (p > 64) { ydyn.resize(p); y = ydyn.data(); }
        for (int k = 0; k < q; ++k) {
            for (int r = 0; r < p; ++r) y[r] = r ? out[(size_t)r * q + k] * tw[(size_t)r * k * tws % n] : out[k];
            if (p == 2) {
                out[k] = y[0] + y[1]; out[q + k] = y[0] - y[1];
            } else if (p == 4) {
                cd a = y[0] + y[2], b = y[0] - y[2], c = y[1] + y[3], d = y[1] - y[3];
                cd jd = (sign < 0) ? cd(d.imag(), -d.real()) : cd(-d.imag(), d.real());
                out[k] = a + c; out[q + k] = b + jd; out[2 * q + k] = a - c; out[3 * q + k] = b - jd;
            } else {
                for (int s = 0; s < p; ++s) {
                    cd acc = y[0];
                    for (int r = 1; r < p; ++r) acc += y[r] * tw[(size_t)((r * s) % p) * q * tws % n];
                    out[(size_t)s * q + k] = acc;
                }
            }
        }
    }
    void exec(const cd* in, int is, cd* out) const { rec(in, is, out, n, 0, 1); }
};

struct FFT2 {
    int n1, n2;
    FFT1 f1, f2;
    std::vector<cd> a, b;
    FFT2(int n1_, int n2_, int sign) : n1(n1_), n2(n2_) {
        f1.init(n1, sign); f2.init(n2, sign);
        a.resize(std::max(n1, n2)); b.resize(std::max(n1, n2));
    }
    // in-place on row-major n1 x n2 data (fftw_plan_dft_2d(n1, n2) convention)
    void exec(cd* d) {
        for (int i = 0; i < n1; ++i) {
            f2.exec(d + (size_t)i * n2, 1, b.data());
            std::memcpy(d + (size_t)i * n2, b.data(), sizeof(cd) * n2);
        }
        for (int j = 0; j < n2; ++j) {
            f1.exec(d + j, n2, b.data());
            for (int i = 0; i < n1; ++i) d[(size_t)i * n2 + j] = b[i];
        }
    }
};

// src/fftw_helper.cpp:36-56 (shift N/2) and :58-79 (shift (N+1)/2).
static void shift2(cd* d, int N1, int N2, int h1, int h2, std::vector<cd>& tmp) {
    tmp.resize((size_t)N1 * N2);
    for (int i = 0; i < N1; ++i)
        for (int j = 0; j < N2; ++j)
            tmp[(size_t)((i + h1) % N1) * N2 + (j + h2) % N2] = d[(size_t)i * N2 + j];
    std::memcpy(d, tmp.data(), sizeof(cd) * N1 * N2);
}
static void fftshift(cd* d, int N1, int N2, std::vector<cd>& t) { shift2(d, N1, N2, N1 / 2, N2 / 2, t); }
static void ifftshift(cd* d, int N1, int N2, std::vector<cd>& t) { shift2(d, N1, N2, (N1 + 1) / 2, (N2 + 1) / 2, t); }

// ------------------------------------------------------------ eigensolver ----
// Cyclic complex Jacobi for an n x n Hermitian matrix (row-major). On return
// U holds eigenvectors in its columns. Any eigenbasis gives the same
// projector sum_{occ} u u^dagger, which is all the initial state needs.
static void jacobi_eigh(int n, const cd* A_in, double* w, cd* U) {
    std::vector<cd> A(A_in, A_in + (size_t)n * n);
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) U[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int i = 0; i < n; ++i) A[i * n + i] = A[i * n + i].real();
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, diag = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) (i == j ? diag : off) += std::norm(A[i * n + j]);
        if (off <= 1e-34 * (diag + 1e-300)) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                cd c = A[p * n + q];
                double ac = std::abs(c);
                if (ac == 0.0) continue;
                double a = A[p * n + p].real(), b = A[q * n + q].real();
                cd ph = c / ac;  // e^{i phi}
                double tau = (b - a) / (2.0 * ac);
                double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                double cs = 1.0 / std::sqrt(1.0 + t * t), sn = t * cs;
                // V = diag(1, conj(ph)) * [[cs, sn], [-sn, cs]]
                cd Vpp = cs, Vpq = sn, Vqp = -std::conj(ph) * sn, Vqq = std::conj(ph) * cs;
                for (int r = 0; r < n; ++r) {  // A <- A V
                    cd x = A[r * n + p], y = A[r * n + q];
                    A[r * n + p] = x * Vpp + y * Vqp;
                    A[r * n + q] = x * Vpq + y * Vqq;
                }
                for (int s = 0; s < n; ++s) {  // A <- V^dagger A
                    cd x = A[p * n + s], y = A[q * n + s];
                    A[p * n + s] = std::conj(Vpp) * x + std::conj(Vqp) * y;
                    A[q * n + s] = std::conj(Vpq) * x + std::conj(Vqq) * y;
                }
                A[p * n + q] = A[q * n + p] = 0.0;
                A[p * n + p] = A[p * n + p].real();
                A[q * n + q] = A[q * n + q].real();
                for (int r = 0; r < n; ++r) {  // U <- U V
                    cd x = U[r * n + p], y = U[r * n + q];
                    U[r * n + p] = x * Vpp + y * Vqp;
                    U[r * n + q] = x * Vpq + y * Vqq;
                }
            }
    }
    for (int i = 0; i < n; ++i) w[i] = A[i * n + i].real();
}

// rho = U diag(occ) U^dagger with occ_n = [Re(U^dagger H U)_nn < 0]
// (src/hamiltonian.cpp:70-101, src/operator.cpp:165-280, src/rdm.cpp:23-38).
static void equilibrium_rho(int nb, const cd* Hk, cd* rho, cd* Hw_back /* may be null */) {
    std::vector<cd> U(nb * nb), HH(nb * nb, 0.0), aux(nb * nb, 0.0);
    std::vector<double> w(nb);
    jacobi_eigh(nb, Hk, w.data(), U.data());
    for (int i = 0; i < nb; ++i)  // aux = U^H Hk
        for (int j = 0; j < nb; ++j) {
            cd s = 0; for (int l = 0; l < nb; ++l) s += std::conj(U[l * nb + i]) * Hk[l * nb + j];
            aux[i * nb + j] = s;
        }
    for (int i = 0; i < nb; ++i)  // HH = aux U
        for (int j = 0; j < nb; ++j) {
            cd s = 0; for (int l = 0; l < nb; ++l) s += aux[i * nb + l] * U[l * nb + j];
            HH[i * nb + j] = s;
        }
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) {
            cd s = 0;
            for (int n = 0; n < nb; ++n)
                if (HH[n * nb + n].real() < 0.0) s += U[i * nb + n] * std::conj(U[j * nb + n]);
            rho[i * nb + j] = s;
        }
    if (Hw_back) {  // U HH U^dagger: the H-gauge round trip of src/swe.cpp:85
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j) {
                cd s = 0; for (int l = 0; l < nb; ++l) s += U[i * nb + l] * HH[l * nb + j];
                aux[i * nb + j] = s;
            }
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j) {
                cd s = 0; for (int l = 0; l < nb; ++l) s += aux[i * nb + l] * std::conj(U[j * nb + l]);
                Hw_back[i * nb + j] = s;
            }
    }
}

// --------------------------------------------------------------- supercell ----
// src/grid.cpp:68-110: n_d = i_d - nr_d/2, idx = i1*nr2 + i2 (nr3 = 1),
// R = n1 a1 + n2 a2 + n3 a3 (Angstrom), later scaled by 1/0.5291772.
struct Cell {
    int N1, N2, N;
    std::vector<int> wann;       // per supercell point: index of the Wannier R placed there, or -1
    std::vector<double> RA;      // N*3, Angstrom
    std::vector<double> Rau;     // N*3, bohr
};

static Cell make_cell(const TB& tb, const Params& p) {
    Cell c; c.N1 = p.nr1; c.N2 = p.nr2; c.N = c.N1 * c.N2;
    c.wann.assign(c.N, -1); c.RA.resize(3 * c.N); c.Rau.resize(3 * c.N);
    for (int i1 = 0; i1 < c.N1; ++i1)
        for (int i2 = 0; i2 < c.N2; ++i2) {
            int n1 = -c.N1 / 2 + i1, n2 = -c.N2 / 2 + i2, n3 = 0;
            int idx = i1 * c.N2 + i2;
            for (int a = 0; a < 3; ++a) {
                double v = n1 * tb.lat[0][a] + n2 * tb.lat[1][a] + n3 * tb.lat[2][a];
                c.RA[3 * idx + a] = v;
                v *= 1.0 / 0.5291772;
                c.Rau[3 * idx + a] = v;
            }
            for (int w = 0; w < tb.nR; ++w)  // first match, as check_vec_in_wannier
                if (tb.R[3 * w] == n1 && tb.R[3 * w + 1] == n2 && tb.R[3 * w + 2] == n3) { c.wann[idx] = w; break; }
        }
    return c;
}

static inline int cell_index(const Cell& c, int n1, int n2) {
    int i1 = ((n1 + c.N1 / 2) % c.N1 + c.N1) % c.N1;
    int i2 = ((n2 + c.N2 / 2) % c.N2 + c.N2) % c.N2;
    return i1 * c.N2 + i2;
}

// Real-space operator fields [N][nb*nb] (SI at placement).
struct Ops {
    int nb;
    std::vector<cd> H;      // eV, later Hartree
    std::vector<cd> X[3];   // rbc = -r, Angstrom, later bohr
    std::vector<cd> V[3];   // eV*Angstrom (never converted: include/constants.h:13)
};

static Ops place_ops(const TB& tb, const Cell& c) {
    Ops o; o.nb = tb.nb;
    const int nb2 = tb.nb * tb.nb;
    o.H.assign((size_t)c.N * nb2, 0.0);
    for (int a = 0; a < 3; ++a) o.X[a].assign((size_t)c.N * nb2, 0.0);
    for (int idx = 0; idx < c.N; ++idx) {
        int w = c.wann[idx];
        if (w < 0) continue;
        for (int e = 0; e < nb2; ++e) {
            o.H[(size_t)idx * nb2 + e] = tb.H[(size_t)w * nb2 + e];
            for (int a = 0; a < 3; ++a) o.X[a][(size_t)idx * nb2 + e] = -tb.r[a][(size_t)w * nb2 + e];
        }
    }
    return o;
}

// Velocity by FFT convolution, literally src/velocity.cpp:41-115 (SI units).
static void velocity_fft(const Cell& c, Ops& o) {
    const int nb = o.nb, nb2 = nb * nb, N = c.N;
    FFT2 fw(c.N1, c.N2, -1), bw(c.N1, c.N2, +1);
    std::vector<cd> hk((size_t)nb2 * N), xk((size_t)nb2 * N), ck((size_t)nb2 * N), t1(N), t2(N), tmp;
    for (int a = 0; a < 3; ++a) {
        for (int e = 0; e < nb2; ++e) {
            for (int r = 0; r < N; ++r) { t1[r] = o.H[(size_t)r * nb2 + e]; t2[r] = o.X[a][(size_t)r * nb2 + e]; }
            fftshift(t1.data(), c.N1, c.N2, tmp); fftshift(t2.data(), c.N1, c.N2, tmp);
            fw.exec(t1.data()); fw.exec(t2.data());
            std::memcpy(&hk[(size_t)e * N], t1.data(), sizeof(cd) * N);
            std::memcpy(&xk[(size_t)e * N], t2.data(), sizeof(cd) * N);
        }
        for (int k = 0; k < N; ++k)
            for (int i = 0; i < nb; ++i)
                for (int j = 0; j < nb; ++j) {
                    cd s = 0;
                    for (int l = 0; l < nb; ++l)
                        s += xk[(size_t)(i * nb + l) * N + k] * hk[(size_t)(l * nb + j) * N + k]
                           - xk[(size_t)(l * nb + j) * N + k] * hk[(size_t)(i * nb + l) * N + k];
                    ck[(size_t)(i * nb + j) * N + k] = s;
                }
        o.V[a].assign((size_t)N * nb2, 0.0);
        for (int e = 0; e < nb2; ++e) {
            for (int k = 0; k < N; ++k) t1[k] = ck[(size_t)e * N + k];
            bw.exec(t1.data());
            for (int k = 0; k < N; ++k) t1[k] = t1[k] / (double)N;
            ifftshift(t1.data(), c.N1, c.N2, tmp);
            for (int r = 0; r < N; ++r) {
                double r0 = c.RA[3 * r + a];
                o.V[a][(size_t)r * nb2 + e] = cd(0.0, 1.0) * (1.0 * r0 * o.H[(size_t)r * nb2 + e] - 1.0 * t1[r]);
            }
        }
    }
}

// The same velocity at one supercell point by direct circular convolution over
// the (sparse) placed blocks: conv(n) = sum_n' rbc(n') H(n-n') - H(n') rbc(n-n').
static void velocity_direct_at(const TB& tb, const Cell& c, const Ops& o, int n1, int n2, int a, cd* out) {
    const int nb = tb.nb, nb2 = nb * nb;
    int idx = cell_index(c, n1, n2);
    std::vector<cd> conv(nb2, 0.0);
    for (int src = 0; src < c.N; ++src) {
        if (c.wann[src] < 0) continue;
        int s1 = -c.N1 / 2 + src / c.N2, s2 = -c.N2 / 2 + src % c.N2;
        int dst = cell_index(c, n1 - s1, n2 - s2);
        if (c.wann[dst] < 0) continue;
        const cd* Xs = &o.X[a][(size_t)src * nb2];
        const cd* Hs = &o.H[(size_t)src * nb2];
        const cd* Xd = &o.X[a][(size_t)dst * nb2];
        const cd* Hd = &o.H[(size_t)dst * nb2];
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j) {
                cd s = 0;
                for (int l = 0; l < nb; ++l) s += Xs[i * nb + l] * Hd[l * nb + j] - Hs[i * nb + l] * Xd[l * nb + j];
                conv[i * nb + j] += s;
            }
    }
    double r0 = c.RA[3 * idx + a];
    for (int e = 0; e < nb2; ++e) out[e] = cd(0.0, 1.0) * (1.0 * r0 * o.H[(size_t)idx * nb2 + e] - 1.0 * conv[e]);
}

// src/swe.cpp:95-111: H * 1.0 / 27.2114, rbc * 1.0 / 0.5291772, v unchanged.
static void ops_to_au(Ops& o) {
    for (auto& h : o.H) h = h * 1.0 / 27.2114;
    for (int a = 0; a < 3; ++a) for (auto& x : o.X[a]) x = x * 1.0 / 0.5291772;
}

// e^{-2 pi i j / N} for j = 0..N-1 (the DFT kernel of FFTW_FORWARD).
static std::vector<cd> dft_table(int N) {
    std::vector<cd> t(N);
    for (int j = 0; j < N; ++j) {
        long double ang = -2.0L * 3.141592653589793238462643383279502884L * j / N;
        t[j] = cd((double)cosl(ang), (double)sinl(ang));
    }
    return t;
}

// ------------------------------------------------------------- fft form ----
// src/swe.cpp:49-93 (setup order) and :183-223 (time loop);
// src/solver.cpp:86-181 (step), :1411-1425 (Heff), :1428-1726 (stages).
static void run_fftform(const TB& tb, const Params& p, double* t_out, double* J_out) {
    Cell c = make_cell(tb, p);
    Ops o = place_ops(tb, c);
    velocity_fft(c, o);            // SI, before the au conversion (src/swe.cpp:66-68 precede :78)
    ops_to_au(o);
    Fields F = make_fields(p);
    const int nb = tb.nb, nb2 = nb * nb, N = c.N, N1 = c.N1, N2 = c.N2, nt = p.nt;
    FFT2 fw(N1, N2, -1), bw(N1, N2, +1);
    std::vector<cd> tmp, t1(N), t2(N);

    // Initial state (src/swe.cpp:80-90): H -> k (ifftshift, FFT), per-k
    // eigenbasis, T=0 occupation, back to W gauge, both to r (IFFT/N, ifftshift).
    std::vector<cd> Hk((size_t)N * nb2), rhok((size_t)N * nb2), Hwk((size_t)N * nb2);
    for (int e = 0; e < nb2; ++e) {
        for (int r = 0; r < N; ++r) t1[r] = o.H[(size_t)r * nb2 + e];
        ifftshift(t1.data(), N1, N2, tmp);
        fw.exec(t1.data());
        for (int k = 0; k < N; ++k) Hk[(size_t)k * nb2 + e] = t1[k];
    }
    for (int k = 0; k < N; ++k)
        equilibrium_rho(nb, &Hk[(size_t)k * nb2], &rhok[(size_t)k * nb2], &Hwk[(size_t)k * nb2]);
    std::vector<cd> rho((size_t)N * nb2);
    for (int e = 0; e < nb2; ++e) {
        for (int k = 0; k < N; ++k) t1[k] = Hwk[(size_t)k * nb2 + e];
        bw.exec(t1.data());
        for (int k = 0; k < N; ++k) t1[k] = t1[k] / (double)N;
        ifftshift(t1.data(), N1, N2, tmp);
        for (int r = 0; r < N; ++r) o.H[(size_t)r * nb2 + e] = t1[r];
        for (int k = 0; k < N; ++k) t1[k] = rhok[(size_t)k * nb2 + e];
        bw.exec(t1.data());
        for (int k = 0; k < N; ++k) t1[k] = t1[k] / (double)N;
        ifftshift(t1.data(), N1, N2, tmp);
        for (int r = 0; r < N; ++r) rho[(size_t)r * nb2 + e] = t1[r];
    }

    std::vector<cd> heff((size_t)N * nb2), K[4], pp(N);
    for (int s = 0; s < 4; ++s) K[s].assign((size_t)N * nb2, 0.0);
    std::vector<cd> hk((size_t)nb2 * N), rk((size_t)nb2 * N), ck((size_t)nb2 * N);
    const double dt = F.dt;

    auto update_heff = [&](const double* e) {
        for (int r = 0; r < N; ++r)
            for (int q = 0; q < nb2; ++q) {
                size_t i = (size_t)r * nb2 + q;
                heff[i] = o.H[i] + e[0] * o.X[0][i] + e[1] * o.X[1][i] + e[2] * o.X[2][i];
            }
    };
    // stage s: K[s] = -i IFFT([FFT(p heff), FFT(rho + c K[s-1])])
    auto stage = [&](int s, double coef) {
        for (int q = 0; q < nb2; ++q) {
            for (int r = 0; r < N; ++r) {
                size_t i = (size_t)r * nb2 + q;
                t1[r] = pp[r] * heff[i];
                t2[r] = (s == 0) ? rho[i] : rho[i] + coef * K[s - 1][i];
            }
            fftshift(t1.data(), N1, N2, tmp); fftshift(t2.data(), N1, N2, tmp);
            fw.exec(t1.data()); fw.exec(t2.data());
            std::memcpy(&hk[(size_t)q * N], t1.data(), sizeof(cd) * N);
            std::memcpy(&rk[(size_t)q * N], t2.data(), sizeof(cd) * N);
        }
        for (int k = 0; k < N; ++k)
            for (int i = 0; i < nb; ++i)
                for (int j = 0; j < nb; ++j) {
                    cd cv = 0;
                    for (int l = 0; l < nb; ++l)
                        cv += hk[(size_t)(i * nb + l) * N + k] * rk[(size_t)(l * nb + j) * N + k]
                            - rk[(size_t)(i * nb + l) * N + k] * hk[(size_t)(l * nb + j) * N + k];
                    ck[(size_t)(i * nb + j) * N + k] = cv;
                }
        for (int q = 0; q < nb2; ++q) {
            for (int k = 0; k < N; ++k) t1[k] = ck[(size_t)q * N + k];
            bw.exec(t1.data());
            for (int k = 0; k < N; ++k) t1[k] = t1[k] / (double)N;
            ifftshift(t1.data(), N1, N2, tmp);
            for (int r = 0; r < N; ++r) K[s][(size_t)r * nb2 + q] = cd(0.0, -1.0) * t1[r];
        }
    };

    for (int ti = 0; ti < nt; ++ti) {
        double ax = F.A[0][ti], ay = F.A[1][ti], az = F.A[2][ti];
        for (int r = 0; r < N; ++r)  // src/swe.cpp:225-233
            pp[r] = std::exp(cd(0.0, +ax * c.Rau[3 * r] + ay * c.Rau[3 * r + 1] + az * c.Rau[3 * r + 2]));
        double e[3][3];
        stage_fields(F, ti, nt, e);
        update_heff(e[0]); stage(0, 0.0);
        update_heff(e[1]); stage(1, 0.5 * dt); stage(2, 0.5 * dt);
        update_heff(e[2]); stage(3, dt);
        cd prefac(0.1666666666666 * dt, 0.0);
        for (size_t i = 0; i < (size_t)N * nb2; ++i)
            rho[i] = rho[i] + prefac * (K[0][i] + 2.0 * K[1][i] + 2.0 * K[2][i] + K[3][i]);
        // src/swe.cpp:197-217: rho *= conj(p); J_x, J_y, J_z; rho *= p
        for (int r = 0; r < N; ++r) for (int q = 0; q < nb2; ++q) rho[(size_t)r * nb2 + q] = rho[(size_t)r * nb2 + q] * std::conj(pp[r]);
        for (int a = 0; a < 3; ++a) {
            cd sum = 0;
            for (int i1 = N1 / 2 - 2; i1 < N1 / 2 + 2; ++i1)
                for (int i2 = N2 / 2 - 2; i2 < N2 / 2 + 2; ++i2)
                    for (int q = 0; q < nb2; ++q) {
                        size_t i = (size_t)(i1 * N2 + i2) * nb2 + q;
                        cd v = o.V[a][i];
                        v *= std::conj(rho[i]);
                        sum += v;
                    }
            J_out[(size_t)ti * 6 + 2 * a] = sum.real();
            J_out[(size_t)ti * 6 + 2 * a + 1] = sum.imag();
        }
        for (int r = 0; r < N; ++r) for (int q = 0; q < nb2; ++q) rho[(size_t)r * nb2 + q] = rho[(size_t)r * nb2 + q] * pp[r];
    }
    for (int ti = 0; ti < nt; ++ti) t_out[ti] = F.t[ti];
}

// --------------------------------------------------------------- per-k form ----
struct PerK {
    int nb, N1, N2, N, nt;
    // Fourier-sum support: supercell points carrying a Wannier block, with
    // their integer (n1, n2), H, X[3] (au); all-zero blocks dropped (exact).
    std::vector<int> sn1, sn2;
    std::vector<cd> sH, sX[3];
    std::vector<double> sR;  // bohr, 3 per point
    // current window: the 16 points n1, n2 in {-2..1}, with v blocks (eV*A)
    std::vector<int> wn1, wn2;
    std::vector<cd> wV[3];
    std::vector<double> wR;
    std::vector<cd> tw1, tw2;
    std::vector<cd> phs, phw;  // e^{i A[ti].R} per step: [nt][ns], [nt][nw]
    Fields F;
};

static bool block_zero(const cd* b, int n) { for (int i = 0; i < n; ++i) if (b[i] != 0.0) return false; return true; }

static PerK setup_perk(const TB& tb, const Params& p) {
    PerK P;
    Cell c = make_cell(tb, p);
    Ops o = place_ops(tb, c);
    const int nb = tb.nb, nb2 = nb * nb;
    P.nb = nb; P.N1 = c.N1; P.N2 = c.N2; P.N = c.N; P.nt = p.nt;
    // window velocity by direct convolution (SI), before the au conversion
    std::vector<cd> vw(nb2);
    for (int n1 = -2; n1 <= 1; ++n1)
        for (int n2 = -2; n2 <= 1; ++n2) {
            int idx = cell_index(c, n1, n2);
            bool any = false;
            std::vector<cd> blk[3];
            for (int a = 0; a < 3; ++a) {
                velocity_direct_at(tb, c, o, n1, n2, a, vw.data());
                blk[a].assign(vw.begin(), vw.end());
                any = any || !block_zero(vw.data(), nb2);
            }
            if (!any) continue;
            P.wn1.push_back(n1); P.wn2.push_back(n2);
            for (int a = 0; a < 3; ++a) { P.wV[a].insert(P.wV[a].end(), blk[a].begin(), blk[a].end()); P.wR.push_back(c.Rau[3 * idx + a]); }
        }
    ops_to_au(o);
    for (int idx = 0; idx < c.N; ++idx) {
        if (c.wann[idx] < 0) continue;
        bool z = block_zero(&o.H[(size_t)idx * nb2], nb2);
        for (int a = 0; a < 3; ++a) z = z && block_zero(&o.X[a][(size_t)idx * nb2], nb2);
        if (z) continue;
        P.sn1.push_back(-c.N1 / 2 + idx / c.N2); P.sn2.push_back(-c.N2 / 2 + idx % c.N2);
        P.sH.insert(P.sH.end(), &o.H[(size_t)idx * nb2], &o.H[(size_t)idx * nb2] + nb2);
        for (int a = 0; a < 3; ++a) {
            P.sX[a].insert(P.sX[a].end(), &o.X[a][(size_t)idx * nb2], &o.X[a][(size_t)idx * nb2] + nb2);
            P.sR.push_back(c.Rau[3 * idx + a]);
        }
    }
    P.tw1 = dft_table(c.N1); P.tw2 = dft_table(c.N2);
    P.F = make_fields(p);
    const int ns = (int)P.sn1.size(), nw = (int)P.wn1.size();
    P.phs.resize((size_t)P.nt * ns); P.phw.resize((size_t)P.nt * nw);
    for (int ti = 0; ti < P.nt; ++ti) {  // src/swe.cpp:225-233
        double ax = P.F.A[0][ti], ay = P.F.A[1][ti], az = P.F.A[2][ti];
        for (int s = 0; s < ns; ++s)
            P.phs[(size_t)ti * ns + s] = std::exp(cd(0.0, +ax * P.sR[3 * s] + ay * P.sR[3 * s + 1] + az * P.sR[3 * s + 2]));
        for (int s = 0; s < nw; ++s)
            P.phw[(size_t)ti * nw + s] = std::exp(cd(0.0, +ax * P.wR[3 * s] + ay * P.wR[3 * s + 1] + az * P.wR[3 * s + 2]));
    }
    return P;
}

static inline int imod(long a, int n) { long r = a % n; return (int)(r < 0 ? r + n : r); }

// H_k = sum_n H(n) e^{-2 pi i m.n / N} over the support (no field, no phase)
static void perk_hk(const PerK& P, int m1, int m2, cd* Hk) {
    const int nb2 = P.nb * P.nb;
    for (int e = 0; e < nb2; ++e) Hk[e] = 0.0;
    for (size_t s = 0; s < P.sn1.size(); ++s) {
        cd w = P.tw1[imod((long)m1 * P.sn1[s], P.N1)] * P.tw2[imod((long)m2 * P.sn2[s], P.N2)];
        for (int e = 0; e < nb2; ++e) Hk[e] += P.sH[s * nb2 + e] * w;
    }
}

// Propagates k-points kidx[0..nk) (k = m1 N2 + m2) over all nt steps; adds
// their current terms to J (nt*3, unnormalised) and, when rho_out is given,
// stores each one's final rho_k at rho_out[i*nb*nb].
static void perk_list(const PerK& P, const int* kidx, int nk, std::vector<cd>& J, cd* rho_out,
                      const cd* rho_init = nullptr /* [N][nb*nb]: start from these states */) {
    const int nb = P.nb, nb2 = nb * nb, nt = P.nt, ns = (int)P.sn1.size(), nw = (int)P.wn1.size();
    std::vector<cd> Hk(nb2), rho(nb2), Ht(nb2), Xt[3], Hs(nb2), Kc[4], rs(nb2), ws(ns), ww(nw), w0s(ns), w0w(nw);
    for (int a = 0; a < 3; ++a) Xt[a].resize(nb2);
    for (int s = 0; s < 4; ++s) Kc[s].resize(nb2);
    const double dt = P.F.dt;
    auto f = [&](const cd* X, cd* out) {  // out = -i [Hs, X]
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j) {
                cd cv = 0;
                for (int l = 0; l < nb; ++l) cv += Hs[i * nb + l] * X[l * nb + j] - X[i * nb + l] * Hs[l * nb + j];
                out[i * nb + j] = cd(0.0, -1.0) * cv;
            }
    };
    auto heff = [&](const double* e) {
        for (int q = 0; q < nb2; ++q) Hs[q] = Ht[q] + e[0] * Xt[0][q] + e[1] * Xt[1][q] + e[2] * Xt[2][q];
    };
    for (int ik = 0; ik < nk; ++ik) {
        const int k = kidx[ik];
        int m1 = k / P.N2, m2 = k % P.N2;
        if (rho_init) {
            std::copy(rho_init + (size_t)k * nb2, rho_init + (size_t)(k + 1) * nb2, rho.begin());
        } else {
            perk_hk(P, m1, m2, Hk.data());
            equilibrium_rho(nb, Hk.data(), rho.data(), nullptr);
        }
        for (int s = 0; s < ns; ++s) w0s[s] = P.tw1[imod((long)m1 * P.sn1[s], P.N1)] * P.tw2[imod((long)m2 * P.sn2[s], P.N2)];
        for (int s = 0; s < nw; ++s) w0w[s] = P.tw1[imod((long)m1 * P.wn1[s], P.N1)] * P.tw2[imod((long)m2 * P.wn2[s], P.N2)];
        for (int ti = 0; ti < nt; ++ti) {
            for (int s = 0; s < ns; ++s) ws[s] = w0s[s] * P.phs[(size_t)ti * ns + s];
            for (int q = 0; q < nb2; ++q) { Ht[q] = 0.0; Xt[0][q] = Xt[1][q] = Xt[2][q] = 0.0; }
            for (int s = 0; s < ns; ++s)
                for (int q = 0; q < nb2; ++q) {
                    Ht[q] += P.sH[s * nb2 + q] * ws[s];
                    for (int a = 0; a < 3; ++a) Xt[a][q] += P.sX[a][s * nb2 + q] * ws[s];
                }
            double e[3][3];
            stage_fields(P.F, ti, nt, e);
            heff(e[0]); f(rho.data(), Kc[0].data());
            heff(e[1]);
            for (int q = 0; q < nb2; ++q) rs[q] = rho[q] + 0.5 * dt * Kc[0][q];
            f(rs.data(), Kc[1].data());
            for (int q = 0; q < nb2; ++q) rs[q] = rho[q] + 0.5 * dt * Kc[1][q];
            f(rs.data(), Kc[2].data());
            heff(e[2]);
            for (int q = 0; q < nb2; ++q) rs[q] = rho[q] + dt * Kc[2][q];
            f(rs.data(), Kc[3].data());
            double pref = 0.1666666666666 * dt;
            for (int q = 0; q < nb2; ++q) rho[q] = rho[q] + pref * (Kc[0][q] + 2.0 * Kc[1][q] + 2.0 * Kc[2][q] + Kc[3][q]);
            for (int s = 0; s < nw; ++s) ww[s] = w0w[s] * P.phw[(size_t)ti * nw + s];
            for (int a = 0; a < 3; ++a) {
                cd sum = 0;
                for (int s = 0; s < nw; ++s) {
                    cd part = 0;
                    for (int q = 0; q < nb2; ++q) part += P.wV[a][s * nb2 + q] * std::conj(rho[q]);
                    sum += part * ww[s];
                }
                J[(size_t)ti * 3 + a] += sum;
            }
        }
        if (rho_out) std::copy(rho.begin(), rho.end(), rho_out + (size_t)ik * nb2);
    }
}

static void perk_range(const PerK& P, int k0, int k1, std::vector<cd>& J /* nt*3 */, const cd* rho_init) {
    std::vector<int> ks;
    for (int k = k0; k < k1; ++k) ks.push_back(k);
    perk_list(P, ks.data(), (int)ks.size(), J, nullptr, rho_init);
}

// rho_init: null = the T = 0 initial state; else [N][nb*nb] states to start
// from (a run that continues a previous one, src/swe.cpp:183-223 called twice)
static void run_perk(const TB& tb, const Params& p, int nthreads, double* t_out, double* J_out,
                     const cd* rho_init = nullptr) {
    PerK P = setup_perk(tb, p);
    if (nthreads < 1) nthreads = 1;
    std::vector<std::vector<cd>> Jp(nthreads, std::vector<cd>((size_t)P.nt * 3, 0.0));
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) {
        int k0 = (int)((long)P.N * t / nthreads), k1 = (int)((long)P.N * (t + 1) / nthreads);
        th.emplace_back(perk_range, std::cref(P), k0, k1, std::ref(Jp[t]), rho_init);
    }
    for (auto& x : th) x.join();
    for (int ti = 0; ti < P.nt; ++ti)
        for (int a = 0; a < 3; ++a) {
            cd s = 0;
            for (int t = 0; t < nthreads; ++t) s += Jp[t][(size_t)ti * 3 + a];
            s = s / (double)P.N;
            J_out[(size_t)ti * 6 + 2 * a] = s.real();
            J_out[(size_t)ti * 6 + 2 * a + 1] = s.imag();
        }
    for (int ti = 0; ti < P.nt; ++ti) t_out[ti] = P.F.t[ti];
}

// ------------------------------------------------------------------ C API ----
#define API extern "C" __attribute__((visibility("default")))
#define GUARD(...) try { __VA_ARGS__; return 0; } catch (const std::exception& ex) { g_err = ex.what(); return 1; }

API const char* oracle_last_error() { return g_err.c_str(); }

API int oracle_nt(const int* nr, const double* dp, int* nt) { GUARD(*nt = make_params(nr, dp).nt) }

API int oracle_model_dims(const char* path, int* nb, int* nR) {
    GUARD(TB tb = parse_tb(path); *nb = tb.nb; *nR = tb.nR)
}

// blocks as parsed (SI): H[nR][nb][nb], r[3][nR][nb][nb] as interleaved complex
API int oracle_model(const char* path, double* lat, int* R, double* H, double* r) {
    GUARD(
        TB tb = parse_tb(path);
        const size_t nb2 = (size_t)tb.nb * tb.nb;
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) lat[3 * i + j] = tb.lat[i][j];
        for (int i = 0; i < 3 * tb.nR; ++i) R[i] = tb.R[i];
        std::memcpy(H, tb.H.data(), sizeof(cd) * nb2 * tb.nR);
        for (int a = 0; a < 3; ++a) std::memcpy(r + 2 * a * nb2 * tb.nR, tb.r[a].data(), sizeof(cd) * nb2 * tb.nR);
    )
}

API int oracle_fields(const int* nr, const double* dp, double* t, double* E, double* A) {
    GUARD(
        Params p = make_params(nr, dp);
        Fields F = make_fields(p);
        for (int i = 0; i < p.nt; ++i) t[i] = F.t[i];
        for (int a = 0; a < 3; ++a) for (int i = 0; i < p.nt; ++i) { E[(size_t)a * p.nt + i] = F.E[a][i]; A[(size_t)a * p.nt + i] = F.A[a][i]; }
    )
}

// H_k (au) on the full k grid, per-k form; out[N][nb*nb] complex
API int oracle_hk(const char* path, const int* nr, const double* dp, double* out) {
    GUARD(
        TB tb = parse_tb(path); Params p = make_params(nr, dp); PerK P = setup_perk(tb, p);
        const int nb2 = P.nb * P.nb;
        for (int k = 0; k < P.N; ++k) perk_hk(P, k / P.N2, k % P.N2, reinterpret_cast<cd*>(out) + (size_t)k * nb2);
    )
}

// initial rho_k; mode 0 = per-k (direct Fourier sum), 1 = FFT form
API int oracle_rho0(const char* path, const int* nr, const double* dp, int mode, double* out) {
    GUARD(
        TB tb = parse_tb(path); Params p = make_params(nr, dp);
        const int nb = tb.nb, nb2 = nb * nb;
        cd* o = reinterpret_cast<cd*>(out);
        if (mode == 0) {
            PerK P = setup_perk(tb, p);
            std::vector<cd> Hk(nb2);
            for (int k = 0; k < P.N; ++k) { perk_hk(P, k / P.N2, k % P.N2, Hk.data()); equilibrium_rho(nb, Hk.data(), o + (size_t)k * nb2, nullptr); }
        } else {
            Cell c = make_cell(tb, p); Ops op = place_ops(tb, c); ops_to_au(op);
            FFT2 fw(c.N1, c.N2, -1);
            std::vector<cd> t1(c.N), tmp, Hk((size_t)c.N * nb2);
            for (int e = 0; e < nb2; ++e) {
                for (int r = 0; r < c.N; ++r) t1[r] = op.H[(size_t)r * nb2 + e];
                ifftshift(t1.data(), c.N1, c.N2, tmp); fw.exec(t1.data());
                for (int k = 0; k < c.N; ++k) Hk[(size_t)k * nb2 + e] = t1[k];
            }
            for (int k = 0; k < c.N; ++k) equilibrium_rho(nb, &Hk[(size_t)k * nb2], o + (size_t)k * nb2, nullptr);
        }
    )
}

// v_a(n) at the 16 window points n1, n2 in {-2..1} (row-major n1 then n2),
// out[3][16][nb*nb]; mode 0 = direct convolution, 1 = FFT (src/velocity.cpp)
API int oracle_velocity_window(const char* path, const int* nr, const double* dp, int mode, double* out) {
    GUARD(
        TB tb = parse_tb(path); Params p = make_params(nr, dp);
        Cell c = make_cell(tb, p); Ops o = place_ops(tb, c);
        const int nb2 = tb.nb * tb.nb;
        cd* ov = reinterpret_cast<cd*>(out);
        if (mode == 1) velocity_fft(c, o);
        for (int a = 0; a < 3; ++a)
            for (int w = 0; w < 16; ++w) {
                int n1 = -2 + w / 4, n2 = -2 + w % 4;
                cd* dst = ov + ((size_t)a * 16 + w) * nb2;
                if (mode == 1) { int idx = cell_index(c, n1, n2); for (int e = 0; e < nb2; ++e) dst[e] = o.V[a][(size_t)idx * nb2 + e]; }
                else velocity_direct_at(tb, c, o, n1, n2, a, dst);
            }
    )
}

// Full solve. mode 0 = per-k (nthreads threads over k), 1 = FFT form.
// t_out[nt] (au), J_out[nt][3] complex as interleaved doubles.
API int oracle_run(const char* path, const int* nr, const double* dp, int mode, int nthreads, double* t_out, double* J_out) {
    GUARD(
        TB tb = parse_tb(path); Params p = make_params(nr, dp);
        if (mode == 0) run_perk(tb, p, nthreads, t_out, J_out);
        else run_fftform(tb, p, t_out, J_out);
    )
}

// Per-k solve that starts from the given states rho_init[N][nb*nb] (complex)
// instead of the initial state: the J(t) of a second run_simulation.
API int oracle_run_from(const char* path, const int* nr, const double* dp, const double* rho_init, int nthreads,
                        double* t_out, double* J_out) {
    GUARD(
        TB tb = parse_tb(path); Params p = make_params(nr, dp);
        run_perk(tb, p, nthreads, t_out, J_out, reinterpret_cast<const cd*>(rho_init));
    )
}

// Final rho_k (after all nt steps, per-k form) of the nk listed k-points,
// k = m1*nr2 + m2; out[nk][nb*nb] complex. nthreads threads over the list.
API int oracle_propagate_k(const char* path, const int* nr, const double* dp, int nk, const int* kidx, int nthreads,
                           double* out) {
    GUARD(
        TB tb = parse_tb(path); Params p = make_params(nr, dp); PerK P = setup_perk(tb, p);
        for (int i = 0; i < nk; ++i)
            if (kidx[i] < 0 || kidx[i] >= P.N) throw std::runtime_error("k index out of range");
        if (nthreads < 1) nthreads = 1;
        const int nb2 = P.nb * P.nb;
        cd* o = reinterpret_cast<cd*>(out);
        std::vector<std::thread> th;
        std::vector<std::vector<cd>> Jp(nthreads, std::vector<cd>((size_t)P.nt * 3, 0.0));
        for (int t = 0; t < nthreads; ++t) {
            const int i0 = (int)((long)nk * t / nthreads), i1 = (int)((long)nk * (t + 1) / nthreads);
            th.emplace_back([&, i0, i1, t] { perk_list(P, kidx + i0, i1 - i0, Jp[t], o + (size_t)i0 * nb2); });
        }
        for (auto& x : th) x.join();
    )
}

// In-place unnormalised 2-D FFT on row-major n1 x n2 complex data.
API int oracle_fft2(double* data, int n1, int n2, int sign) {
    GUARD(FFT2 f(n1, n2, sign); f.exec(reinterpret_cast<cd*>(data)))
}

API int oracle_eigh(int n, const double* A, double* w, double* U) {
    GUARD(jacobi_eigh(n, reinterpret_cast<const cd*>(A), w, reinterpret_cast<cd*>(U)))
}

API int oracle_equilibrium_rho(int nb, const double* Hk, double* rho) {
    GUARD(equilibrium_rho(nb, reinterpret_cast<const cd*>(Hk), reinterpret_cast<cd*>(rho), nullptr))
}
