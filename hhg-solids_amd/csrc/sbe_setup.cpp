// sbe_setup.cpp -- host setup of one k-space solve, in FP64.
//
// Builds from a _tb.dat model and the Settings block what the propagation
// kernels consume: the Fourier-sum support with its blocks, the per-step
// table of stage fields and Peierls phases, the DFT phase tables and the
// initial density matrices. Compiled with -ffp-contract=off: the field tables
// must be bit-identical to a strict-IEEE evaluation of src/efield.cpp:30-50.
//
// Reference semantics followed (file:line in HHG-Solids):
//   parser            src/wannier_tb.cpp:15-113
//   supercell/grids   src/grid.cpp:64-110, 113-137
//   placement         src/hamiltonian.cpp:8-39, src/berry_connection.cpp:7-50
//                     (exact integer R; the reference truncates a Cramer solve)
//   velocity          src/velocity.cpp:41-115 (direct circular convolution here)
//   unit conversion   src/swe.cpp:95-111, include/constants.h:6-16
//   fields            src/efield.cpp:30-50, 66-77
//   stage fields      src/solver.cpp:109-148
//   Peierls phase     src/swe.cpp:225-233
//   initial state     src/swe.cpp:80-90, src/hamiltonian.cpp:70-101,
//                     src/operator.cpp:165-280, src/rdm.cpp:23-38
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <thread>

#include "sbe_host.h"

namespace hhg {

TBModel read_tb(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open Wannier tb file '" + path + "'");
    TBModel m;
    std::string ln;
    auto next = [&]() -> std::string& {
        if (!std::getline(in, ln)) throw std::runtime_error("truncated tb file '" + path + "'");
        return ln;
    };
    m.comment = next();
    for (auto& row : m.lat) {
        std::istringstream is(next());
        if (!(is >> row[0] >> row[1] >> row[2])) throw std::runtime_error("bad lattice row in tb file");
    }
    m.nb = std::stoi(next());
    m.nR = std::stoi(next());
    if (m.nb < 1 || m.nb > 64 || m.nR < 1) throw std::runtime_error("tb file: bad orbital or R count");
    for (int l = 0; l < (m.nR - 1) / 15 + 1; ++l) {
        std::istringstream is(next());
        int d;
        while (is >> d) m.deg.push_back(d);
    }
    next();
    const int nb = m.nb, nb2 = nb * nb;
    m.R.resize(3 * m.nR);
    m.H.assign((size_t)m.nR * nb2, 0.0);
    for (auto& a : m.r) a.assign((size_t)m.nR * nb2, 0.0);
    for (int i = 0; i < m.nR; ++i) {
        std::istringstream rs(next());
        if (!(rs >> m.R[3 * i] >> m.R[3 * i + 1] >> m.R[3 * i + 2])) throw std::runtime_error("tb file: bad R line");
        for (int j = 0; j < nb2; ++j) {
            std::istringstream ls(next());
            int a, b; double re, im;
            if (!(ls >> a >> b >> re >> im) || a < 1 || a > nb || b < 1 || b > nb)
                throw std::runtime_error("tb file: bad hopping line");
            m.H[(size_t)i * nb2 + (b - 1) * nb + (a - 1)] = cplx(re, im);
        }
        next();
    }
    for (int i = 0; i < m.nR; ++i) {
        next();
        for (int j = 0; j < nb2; ++j) {
            std::istringstream ls(next());
            int a, b; double v[6];
            if (!(ls >> a >> b >> v[0] >> v[1] >> v[2] >> v[3] >> v[4] >> v[5]) || a < 1 || a > nb || b < 1 || b > nb)
                throw std::runtime_error("tb file: bad position line");
            for (int x = 0; x < 3; ++x) m.r[x][(size_t)i * nb2 + (b - 1) * nb + (a - 1)] = cplx(v[2 * x], v[2 * x + 1]);
        }
        if (i + 1 < m.nR) next();
    }
    return m;
}

namespace {

constexpr long double kTwoPi = 6.283185307179586476925286766559005768L;

int wrap_centered(int n, int N) {  // map n to [-N/2, N/2-1] modulo N
    int i = ((n + N / 2) % N + N) % N;
    return i - N / 2;
}

}  // namespace

// Hermitian eigenproblem by cyclic Jacobi rotations; Q gets the eigenvectors
// as columns. Only the occupied-subspace projector is used downstream, which
// does not depend on the eigenvector phases or on degenerate mixing.
void herm_jacobi(int n, std::vector<cplx>& a, std::vector<cplx>& q) {
    q.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) q[(size_t)i * n + i] = 1.0;
    for (int it = 0; it < 60; ++it) {
        double off = 0, on = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) (i == j ? on : off) += std::norm(a[(size_t)i * n + j]);
        if (off <= 1e-32 * on || off == 0.0) return;
        for (int p = 0; p < n - 1; ++p)
            for (int r = p + 1; r < n; ++r) {
                const cplx apr = a[(size_t)p * n + r];
                const double g = std::abs(apr);
                if (g == 0.0) continue;
                const cplx u = apr / g;
                const double app = a[(size_t)p * n + p].real(), arr = a[(size_t)r * n + r].real();
                const double th = 0.5 * (arr - app) / g;
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                // unitary G acting on columns p, r: [c, s; -s*conj(u), c*conj(u)]
                const cplx g00 = c, g01 = s, g10 = -s * std::conj(u), g11 = c * std::conj(u);
                for (int i = 0; i < n; ++i) {
                    cplx x = a[(size_t)i * n + p], y = a[(size_t)i * n + r];
                    a[(size_t)i * n + p] = x * g00 + y * g10;
                    a[(size_t)i * n + r] = x * g01 + y * g11;
                    x = q[(size_t)i * n + p]; y = q[(size_t)i * n + r];
                    q[(size_t)i * n + p] = x * g00 + y * g10;
                    q[(size_t)i * n + r] = x * g01 + y * g11;
                }
                for (int j = 0; j < n; ++j) {
                    cplx x = a[(size_t)p * n + j], y = a[(size_t)r * n + j];
                    a[(size_t)p * n + j] = std::conj(g00) * x + std::conj(g10) * y;
                    a[(size_t)r * n + j] = std::conj(g01) * x + std::conj(g11) * y;
                }
                a[(size_t)p * n + r] = 0.0; a[(size_t)r * n + p] = 0.0;
                a[(size_t)p * n + p] = a[(size_t)p * n + p].real();
                a[(size_t)r * n + r] = a[(size_t)r * n + r].real();
            }
    }
}

namespace {

// rho = Q diag(occ) Q^dagger with occ_n = [Re(Q^dagger H Q)_nn < 0]
void t0_projector(int nb, const cplx* hk, cplx* rho) {
    std::vector<cplx> a(hk, hk + (size_t)nb * nb), q;
    herm_jacobi(nb, a, q);
    std::vector<char> occ(nb);
    for (int n = 0; n < nb; ++n) {  // diagonal of Q^dagger H Q from the original H
        cplx d = 0;
        for (int i = 0; i < nb; ++i) {
            cplx hq = 0;
            for (int j = 0; j < nb; ++j) hq += hk[(size_t)i * nb + j] * q[(size_t)j * nb + n];
            d += std::conj(q[(size_t)i * nb + n]) * hq;
        }
        occ[n] = d.real() < 0.0;
    }
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) {
            cplx s = 0;
            for (int n = 0; n < nb; ++n)
                if (occ[n]) s += q[(size_t)i * nb + n] * std::conj(q[(size_t)j * nb + n]);
            rho[(size_t)i * nb + j] = s;
        }
}

}  // namespace

HostSolve build_solve(const TBModel& m, const SolveParams& p, bool host_state) {
    if (p.nr3 != 1) throw std::runtime_error("nr3 must be 1: the reference propagates 2-D supercells only");
    if (p.nr1 < 4 || p.nr2 < 4 || (p.nr1 % 2) || (p.nr2 % 2))
        throw std::runtime_error("nr1 and nr2 must be even and >= 4 (odd grids break the reference's fftshift pair)");
    const int nt = (int)(p.tmax / p.dt);
    if (nt < 2) throw std::runtime_error("tmax/dt must give at least 2 time steps");
    const int nb = m.nb, nb2 = nb * nb, N1 = p.nr1, N2 = p.nr2, N = N1 * N2;

    HostSolve hs;
    SolveDesc& d = hs.d;
    std::memset(&d, 0, sizeof(d));
    d.nb = nb; d.N1 = N1; d.N2 = N2; d.N = N; d.nt = nt;

    // Wannier R vectors that lie on the supercell n_d in [-N_d/2, N_d/2-1];
    // the first occurrence wins (src/wannier_tb.cpp:219-228).
    struct Pt { int n1, n2, w; };
    std::vector<Pt> placed;
    for (int w = 0; w < m.nR; ++w) {
        int a = m.R[3 * w], b = m.R[3 * w + 1], c = m.R[3 * w + 2];
        if (c != 0 || a < -N1 / 2 || a > N1 / 2 - 1 || b < -N2 / 2 || b > N2 / 2 - 1) continue;
        bool dup = false;
        for (auto& q : placed) dup = dup || (q.n1 == a && q.n2 == b);
        if (!dup) placed.push_back({a, b, w});
    }
    auto block_at = [&](int n1, int n2) -> int {
        for (auto& q : placed) if (q.n1 == n1 && q.n2 == n2) return q.w;
        return -1;
    };
    auto RA = [&](int n1, int n2, int ax) {  // src/grid.cpp:81-93 (Angstrom)
        return n1 * m.lat[0][ax] + n2 * m.lat[1][ax] + 0 * m.lat[2][ax];
    };

    // v_a on the current window n1, n2 in {-2..1} (src/observable.cpp:18-19),
    // v = i (R_a H(R) - sum_R' [rbc(R') H(R-R') - H(R') rbc(R-R')]), SI units,
    // circular on the supercell (src/velocity.cpp:41-115).
    std::vector<cplx> vwin(3 * 16 * (size_t)nb2, 0.0);
    for (auto& a : placed)
        for (auto& b : placed) {
            int t1 = wrap_centered(a.n1 + b.n1, N1), t2 = wrap_centered(a.n2 + b.n2, N2);
            if (t1 < -2 || t1 > 1 || t2 < -2 || t2 > 1) continue;
            int wdx = (t1 + 2) * 4 + (t2 + 2);
            for (int ax = 0; ax < 3; ++ax) {
                const cplx* Xa = &m.r[ax][(size_t)a.w * nb2];  // rbc = -r
                const cplx* Hb = &m.H[(size_t)b.w * nb2];
                const cplx* Ha = &m.H[(size_t)a.w * nb2];
                const cplx* Xb = &m.r[ax][(size_t)b.w * nb2];
                cplx* out = &vwin[((size_t)ax * 16 + wdx) * nb2];
                for (int i = 0; i < nb; ++i)
                    for (int j = 0; j < nb; ++j) {
                        cplx s = 0;
                        for (int l = 0; l < nb; ++l)
                            s += (-Xa[i * nb + l]) * Hb[l * nb + j] - Ha[i * nb + l] * (-Xb[l * nb + j]);
                        out[i * nb + j] += s;  // accumulates the convolution
                    }
            }
        }
    for (int t1 = -2; t1 <= 1; ++t1)
        for (int t2 = -2; t2 <= 1; ++t2) {
            int wdx = (t1 + 2) * 4 + (t2 + 2), w = block_at(t1, t2);
            for (int ax = 0; ax < 3; ++ax) {
                cplx* v = &vwin[((size_t)ax * 16 + wdx) * nb2];
                double r0 = RA(t1, t2, ax);
                for (int e = 0; e < nb2; ++e) {
                    cplx h = (w >= 0) ? m.H[(size_t)w * nb2 + e] : cplx(0.0);
                    v[e] = cplx(0.0, 1.0) * (1.0 * r0 * h - 1.0 * v[e]);
                }
            }
        }

    // Field tables (src/efield.cpp:30-50 in au, src/grid.cpp:64-67,133-135).
    hs.t.resize(nt);
    for (int i = 0; i < nt; ++i) { hs.t[i] = (double)i * p.dt; hs.t[i] *= 1.0 / 0.02418884; }
    double omega = 2.0 * M_PI * 2997.025 / (p.lambda_nm * 10.0);
    omega *= 1.0 / (1.0 / 0.02418884);
    double tmf = p.tmax_field;
    tmf *= 1.0 / 0.02418884;
    const double e0 = sqrtf(p.intensity * 1.0 / 3.56e16);
    const double dt = hs.t[1] - hs.t[0];
    hs.dt = dt;
    for (int ax = 0; ax < 3; ++ax) { hs.E[ax].assign(nt, 0.0); hs.A[ax].assign(nt, 0.0); }
    for (int i = 0; i < nt; ++i) {
        double env = 0.0;
        if (hs.t[i] < tmf) { double sn = sin(M_PI * hs.t[i] / tmf); env = sn * sn; }
        for (int ax = 0; ax < 3; ++ax) hs.E[ax][i] = p.s[ax] * e0 * env * sinf(omega * hs.t[i] + p.phi[ax]);
    }
    for (int i = 1; i < nt; ++i)
        for (int ax = 0; ax < 3; ++ax) hs.A[ax][i] = hs.A[ax][i - 1] - hs.E[ax][i - 1] * dt;
    for (int ax = 0; ax < 3; ++ax) {
        bool live = false;
        for (int i = 0; i < nt; ++i) live = live || hs.E[ax][i] != 0.0;
        if (live) d.xmask |= 1 << ax;
    }

    // Support: union of the placed points with a non-zero H or live rbc
    // block and the window points with a non-zero v block.
    struct Sup { int n1, n2; std::vector<cplx> b; uint32_t om; };
    std::vector<Sup> sup;
    auto find_sup = [&](int n1, int n2) -> Sup& {
        for (auto& s : sup) if (s.n1 == n1 && s.n2 == n2) return s;
        sup.push_back({n1, n2, std::vector<cplx>((size_t)NOPS * nb2, 0.0), 0u});
        return sup.back();
    };
    auto nonzero = [&](const cplx* b) { for (int e = 0; e < nb2; ++e) if (b[e] != 0.0) return true; return false; };
    for (auto& q : placed) {
        const cplx* H = &m.H[(size_t)q.w * nb2];
        bool hz = nonzero(H), xz[3];
        for (int ax = 0; ax < 3; ++ax) xz[ax] = ((d.xmask >> ax) & 1) && nonzero(&m.r[ax][(size_t)q.w * nb2]);
        if (!hz && !xz[0] && !xz[1] && !xz[2]) continue;
        Sup& s = find_sup(q.n1, q.n2);
        if (hz) {
            s.om |= 1u << OP_H;
            for (int e = 0; e < nb2; ++e) s.b[(size_t)OP_H * nb2 + e] = H[e] * 1.0 / 27.2114;
        }
        for (int ax = 0; ax < 3; ++ax)
            if (xz[ax]) {
                s.om |= 1u << (OP_X + ax);
                for (int e = 0; e < nb2; ++e)
                    s.b[(size_t)(OP_X + ax) * nb2 + e] = (-m.r[ax][(size_t)q.w * nb2 + e]) * 1.0 / 0.5291772;
            }
    }
    for (int t1 = -2; t1 <= 1; ++t1)
        for (int t2 = -2; t2 <= 1; ++t2) {
            int wdx = (t1 + 2) * 4 + (t2 + 2);
            for (int ax = 0; ax < 3; ++ax) {
                const cplx* v = &vwin[((size_t)ax * 16 + wdx) * nb2];
                if (!nonzero(v)) continue;
                Sup& s = find_sup(t1, t2);
                s.om |= 1u << (OP_V + ax);
                d.vmask |= 1 << ax;
                for (int e = 0; e < nb2; ++e) s.b[(size_t)(OP_V + ax) * nb2 + e] = v[e];
            }
        }
    if ((int)sup.size() > MAX_SUPPORT) throw std::runtime_error("model support exceeds MAX_SUPPORT points");
    // rbc axes only count as live when some block is non-zero
    {
        int xm = 0;
        for (auto& q : sup)
            for (int ax = 0; ax < 3; ++ax)
                if (q.om & (1u << (OP_X + ax))) xm |= 1 << ax;
        d.xmask &= xm;
    }
    // Hermitian pairing: B(-n) = B(n)^dagger (to 1e-13 of the block scale,
    // which covers the rounding of the convolved velocity), checked for the
    // H and rbc blocks and, separately, for the v blocks of the window.
    auto pair_ok = [&](int o_lo, int o_hi) {
        for (size_t i = 0; i < sup.size(); ++i) {
            const Sup& a = sup[i];
            const Sup* b = nullptr;
            for (auto& q : sup) if (q.n1 == -a.n1 && q.n2 == -a.n2) b = &q;
            for (int o = o_lo; o < o_hi; ++o) {
                const cplx* A = &a.b[(size_t)o * nb2];
                double scale = 0.0;
                for (int e = 0; e < nb2; ++e) scale = std::max(scale, std::abs(A[e]));
                if (scale == 0.0 && !(b && (b->om & (1u << o)))) continue;
                if (!b) return false;
                const cplx* B = &b->b[(size_t)o * nb2];
                for (int e = 0; e < nb2; ++e) scale = std::max(scale, std::abs(B[e]));
                for (int i1 = 0; i1 < nb; ++i1)
                    for (int j1 = 0; j1 < nb; ++j1)
                        if (std::abs(B[j1 * nb + i1] - std::conj(A[i1 * nb + j1])) > 1e-13 * scale) return false;
            }
        }
        return true;
    };
    const bool herm_hx = pair_ok(OP_H, OP_V);
    const bool vpaired = pair_ok(OP_V, NOPS);
    d.herm = herm_hx && vpaired;
    const int ns = (int)sup.size();
    d.ns = ns;
    d.stride = STEP_HEAD + 2 * ns;
    hs.blk.assign((size_t)ns * NOPS * nb2 * 2, 0.0);
    for (int s = 0; s < ns; ++s) {
        d.n1[s] = sup[s].n1; d.n2[s] = sup[s].n2; d.opmask[s] = sup[s].om;
        const bool origin = sup[s].n1 == 0 && sup[s].n2 == 0;
        const bool rep = sup[s].n1 > 0 || (sup[s].n1 == 0 && sup[s].n2 > 0);
        d.role[s] = origin ? ROLE_ORIGIN : (rep ? ROLE_REP : ROLE_PARTNER);
        std::memcpy(&hs.blk[(size_t)s * NOPS * nb2 * 2], sup[s].b.data(), sizeof(cplx) * NOPS * nb2);
    }

    for (int ax = 0; ax < 3; ++ax) { hs.lat[ax] = m.lat[0][ax]; hs.lat[3 + ax] = m.lat[1][ax]; }

    // Supports grouped by n2 for the row-separable kernels.
    d.hermhx = herm_hx ? 1 : 0;
    {
        std::vector<int> vals;
        for (int s = 0; s < ns; ++s)
            if (std::find(vals.begin(), vals.end(), d.n2[s]) == vals.end()) vals.push_back(d.n2[s]);
        std::sort(vals.begin(), vals.end());
        d.nd = 0;
        d.nwd = 0;
        if ((int)vals.size() <= SEP_MAXD) {
            d.nd = (int)vals.size();
            int q = 0;
            for (int dd = 0; dd < d.nd; ++dd) {
                d.dn2[dd] = vals[dd];
                d.dseg[dd] = q;
                bool hasv = false;
                for (int s = 0; s < ns; ++s)
                    if (d.n2[s] == vals[dd]) {
                        d.sorder[q++] = s;
                        hasv = hasv || (d.opmask[s] & (7u << OP_V));
                    }
                d.wdi[dd] = hasv ? d.nwd++ : -1;
            }
            d.dseg[d.nd] = q;
        }
    }

    // Per-step table: stage fields and the Peierls phase of every support point.
    std::vector<double> Rb(3 * (size_t)ns);
    for (int s = 0; s < ns; ++s)
        for (int ax = 0; ax < 3; ++ax) { double v = RA(sup[s].n1, sup[s].n2, ax); v *= 1.0 / 0.5291772; Rb[3 * s + ax] = v; }
    hs.steptab.assign((size_t)nt * d.stride, 0.0);
    for (int ti = 0; ti < nt; ++ti) {
        double* row = &hs.steptab[(size_t)ti * d.stride];
        for (int ax = 0; ax < 3; ++ax) {
            if (ti < nt - 1) {
                row[ax] = hs.E[ax][ti];
                row[3 + ax] = 0.5 * hs.E[ax][ti] + 0.5 * hs.E[ax][ti + 1];
                row[6 + ax] = hs.E[ax][ti + 1];
            } else {
                row[ax] = row[3 + ax] = row[6 + ax] = hs.E[ax][nt - 1];
            }
        }
        const double ax_ = hs.A[0][ti], ay_ = hs.A[1][ti], az_ = hs.A[2][ti];
        for (int s = 0; s < ns; ++s) {
            cplx ph = std::exp(cplx(0.0, +ax_ * Rb[3 * s] + ay_ * Rb[3 * s + 1] + az_ * Rb[3 * s + 2]));
            row[STEP_HEAD + 2 * s] = ph.real();
            row[STEP_HEAD + 2 * s + 1] = ph.imag();
        }
    }

    // Pair packs of the Hermitian nb = 2 kernels. Pairs are represented by
    // their canonical point (n1 > 0, or n1 == 0 and n2 > 0).
    hs.pd = PairDesc{};
    if (nb == 2 && herm_hx) {
        PairDesc& pd = hs.pd;
        for (int ax = 0; ax < 3; ++ax) {
            if (d.xmask & (1 << ax)) pd.xaxis[pd.nx++] = ax;
            if (d.vmask & (1 << ax)) pd.vaxis[pd.nv++] = ax;
        }
        std::vector<std::pair<int, int>> reps;
        for (auto& q : sup) {
            if (q.n1 == 0 && q.n2 == 0) continue;
            const bool canon = q.n1 > 0 || (q.n1 == 0 && q.n2 > 0);
            const std::pair<int, int> c = canon ? std::make_pair(q.n1, q.n2) : std::make_pair(-q.n1, -q.n2);
            if (std::find(reps.begin(), reps.end(), c) == reps.end()) reps.push_back(c);
        }
        if ((int)reps.size() <= PAIR_MAX && pd.nv > 0) {
            pd.nrep = (int)reps.size();
            pd.vpaired = vpaired ? 1 : 0;
            auto at = [&](int n1, int n2) -> const Sup* {
                for (auto& q : sup) if (q.n1 == n1 && q.n2 == n2) return &q;
                return nullptr;
            };
            static const cplx zero4[4] = {0.0, 0.0, 0.0, 0.0};
            auto blk_at = [&](const Sup* q, int o) -> const cplx* { return q ? &q->b[(size_t)o * 4] : zero4; };
            // H / rbc pair coefficients: 4 real outputs (h00, h11, Re h01, Im h01)
            auto pair_hx = [&](const cplx* B, double* out) {
                const cplx b00 = B[0], b01 = B[1], b10 = B[2], b11 = B[3];
                out[0] = 2 * b00.real(); out[1] = -2 * b00.imag();
                out[2] = 2 * b11.real(); out[3] = -2 * b11.imag();
                out[4] = b01.real() + b10.real(); out[5] = -b01.imag() - b10.imag();
                out[6] = b01.imag() - b10.imag(); out[7] = b01.real() - b10.real();
            };
            auto origin_hx = [&](const cplx* B, double* out) {
                out[0] = B[0].real(); out[1] = B[3].real(); out[2] = B[1].real(); out[3] = B[1].imag();
            };
            // current: u(rho) = L . (p, q, Re c, Im c), L = (V00, V11, V01 + V10, i (V10 - V01))
            auto Lvec = [](const cplx* V, cplx* L) {
                L[0] = V[0]; L[1] = V[3]; L[2] = V[1] + V[2]; L[3] = cplx(0.0, 1.0) * (V[2] - V[1]);
            };
            const Sup* org = at(0, 0);
            // streamed pack
            hs.scoef.assign(PS_HEAD + (size_t)PAIR_MAX * PS_REP, 0.0);
            double* sc = hs.scoef.data();
            origin_hx(blk_at(org, OP_H), sc);
            for (int x = 0; x < pd.nx; ++x) origin_hx(blk_at(org, OP_X + pd.xaxis[x]), sc + 4 + 4 * x);
            for (int v = 0; v < pd.nv; ++v) {
                cplx L[4];
                Lvec(blk_at(org, OP_V + pd.vaxis[v]), L);
                for (int j = 0; j < 4; ++j) { sc[16 + 8 * v + j] = L[j].real(); sc[16 + 8 * v + 4 + j] = L[j].imag(); }
            }
            for (int r = 0; r < pd.nrep; ++r) {
                const int n1 = reps[r].first, n2 = reps[r].second;
                pd.n1[r] = n1; pd.n2[r] = n2;
                const Sup* P = at(n1, n2);
                const Sup* M = at(-n1, -n2);
                double* cr = sc + PS_HEAD + PS_REP * r;
                // H and rbc from the representative (the partner holds its dagger)
                const Sup* src = P ? P : M;
                const bool from_partner = (P == nullptr);
                auto hx_block = [&](int o, cplx* B) {
                    const cplx* b = blk_at(src, o);
                    if (!from_partner) { for (int e = 0; e < 4; ++e) B[e] = b[e]; return; }
                    B[0] = std::conj(b[0]); B[1] = std::conj(b[2]); B[2] = std::conj(b[1]); B[3] = std::conj(b[3]);
                };
                cplx B[4];
                hx_block(OP_H, B);
                pair_hx(B, cr);
                for (int x = 0; x < pd.nx; ++x) { hx_block(OP_X + pd.xaxis[x], B); pair_hx(B, cr + 8 + 8 * x); }
                for (int v = 0; v < pd.nv; ++v) {
                    cplx L[4], Lp[4];
                    Lvec(blk_at(P, OP_V + pd.vaxis[v]), L);    // point +n: weight w
                    Lvec(blk_at(M, OP_V + pd.vaxis[v]), Lp);   // point -n: weight conj(w)
                    double* cv = cr + 32 + 16 * v;
                    for (int j = 0; j < 4; ++j) {
                        cv[2 * j] = L[j].real() + Lp[j].real();       // Re J, * Re w
                        cv[2 * j + 1] = -L[j].imag() + Lp[j].imag();  // Re J, * Im w
                        cv[8 + 2 * j] = L[j].imag() + Lp[j].imag();   // Im J, * Re w
                        cv[8 + 2 * j + 1] = L[j].real() - Lp[j].real();  // Im J, * Im w
                    }
                }
            }
            // register pack (real current) when the v blocks pair as well
            if (vpaired) {
                hs.pcoef.assign(PAIR_HEAD + (size_t)PAIR_MAX * PAIR_REP, 0.0);
                double* c = hs.pcoef.data();
                for (int j = 0; j < 16; ++j) c[j] = sc[j];
                for (int v = 0; v < pd.nv; ++v)  // v origin: (V00, V11, 2 Re V01, 2 Im V01) = Re L for Hermitian V(0)
                    for (int j = 0; j < 4; ++j) c[16 + 4 * v + j] = sc[16 + 8 * v + j];
                for (int r = 0; r < pd.nrep; ++r) {
                    const double* cr = sc + PS_HEAD + PS_REP * r;
                    double* o = c + PAIR_HEAD + PAIR_REP * r;
                    for (int j = 0; j < 32; ++j) o[j] = cr[j];
                    for (int v = 0; v < pd.nv; ++v)
                        for (int j = 0; j < 8; ++j) o[32 + 8 * v + j] = cr[32 + 16 * v + j];
                }
            }
            // per-step rows: stage fields of the rbc slots and e^{i A.R} of the pair points
            pd.stride = PAIR_STEP_HEAD + 2 * PAIR_MAX;
            hs.psteps.assign((size_t)nt * pd.stride, 0.0);
            std::vector<double> Rr(3 * (size_t)pd.nrep);
            for (int r = 0; r < pd.nrep; ++r)
                for (int ax = 0; ax < 3; ++ax) {
                    double v = RA(pd.n1[r], pd.n2[r], ax);
                    v *= 1.0 / 0.5291772;
                    Rr[3 * r + ax] = v;
                }
            for (int ti = 0; ti < nt; ++ti) {
                const double* row = &hs.steptab[(size_t)ti * d.stride];
                double* out = &hs.psteps[(size_t)ti * pd.stride];
                for (int j = 0; j < 3; ++j)
                    for (int x = 0; x < pd.nx; ++x) out[3 * j + x] = row[3 * j + pd.xaxis[x]];
                const double ax_ = hs.A[0][ti], ay_ = hs.A[1][ti], az_ = hs.A[2][ti];
                for (int r = 0; r < pd.nrep; ++r) {
                    const cplx ph = std::exp(cplx(0.0, +ax_ * Rr[3 * r] + ay_ * Rr[3 * r + 1] + az_ * Rr[3 * r + 2]));
                    out[PAIR_STEP_HEAD + 2 * r] = ph.real();
                    out[PAIR_STEP_HEAD + 2 * r + 1] = ph.imag();
                }
            }
        } else {
            pd = PairDesc{};
        }
    }

    // DFT phases e^{-2 pi i j/N} (FFTW_FORWARD kernel).
    auto table = [](int n, std::vector<double>& out) {
        out.resize(2 * (size_t)n);
        for (int j = 0; j < n; ++j) {
            long double a = -kTwoPi * j / n;
            out[2 * j] = (double)cosl(a); out[2 * j + 1] = (double)sinl(a);
        }
    };
    table(N1, hs.tw1); table(N2, hs.tw2);

    // Initial state: per k, H_k = sum_s H_s e^{-2 pi i m.n_s/N} and the T=0
    // projector on its negative-energy bands, element-major on the device.
    // Solves bound for the device may leave it to k_init_state (rho0 empty).
    if (!host_state) return hs;
    hs.rho0.assign((size_t)2 * nb2 * N, 0.0);
    auto work = [&](int k0, int k1) {
        std::vector<cplx> hk(nb2), rho(nb2);
        for (int k = k0; k < k1; ++k) {
            int m1 = k / N2, m2 = k % N2;
            std::fill(hk.begin(), hk.end(), 0.0);
            for (int s = 0; s < ns; ++s) {
                if (!(d.opmask[s] & (1u << OP_H))) continue;
                int j1 = (int)(((long)m1 * d.n1[s] % N1 + N1) % N1), j2 = (int)(((long)m2 * d.n2[s] % N2 + N2) % N2);
                cplx w = cplx(hs.tw1[2 * j1], hs.tw1[2 * j1 + 1]) * cplx(hs.tw2[2 * j2], hs.tw2[2 * j2 + 1]);
                for (int e = 0; e < nb2; ++e) hk[e] += sup[s].b[(size_t)OP_H * nb2 + e] * w;
            }
            t0_projector(nb, hk.data(), rho.data());
            for (int e = 0; e < nb2; ++e) {
                hs.rho0[2 * ((size_t)e * N + k)] = rho[e].real();
                hs.rho0[2 * ((size_t)e * N + k) + 1] = rho[e].imag();
            }
        }
    };
    int nth = (int)std::thread::hardware_concurrency();
    nth = std::max(1, std::min(nth, std::min(32, N / 256 + 1)));
    std::vector<std::thread> pool;
    for (int t = 0; t < nth; ++t) pool.emplace_back(work, (int)((long)N * t / nth), (int)((long)N * (t + 1) / nth));
    for (auto& th : pool) th.join();
    return hs;
}

}  // namespace hhg
