// sbe_types.h -- data shared between the host setup and the HIP kernels.
//
// One solve ("instance") is an nr1 x nr2 k-grid of independent nb x nb RK4
// systems. Device layout in HBM, all complex values interleaved (re, im):
//
//   blk      [ns][NOPS][nb*nb]     Wannier blocks at the support points:
//                                  op 0 = H (Hartree), 1..3 = rbc_x,y,z (bohr),
//                                  4..6 = v_x,y,z (eV*Angstrom, window only)
//   steptab  [nt][stride] doubles  per step: 9 stage fields (E1, E2, E3 as
//                                  x,y,z), 3 pad, then e^{i A(t_i).R_s} per
//                                  support point s (ns complex)
//   tw1/tw2  [N1], [N2] complex    e^{-2 pi i j / N}
//   rho      [nb*nb][N] complex    element-major, k fastest (coalesced)
//   jpart    [nt][nwaves][6]       per-wave partial current (re/im of x,y,z)
//   J        [nt][6]               reduced current, normalised by 1/N
#pragma once
#include <cstdint>

namespace hhg {

enum { OP_H = 0, OP_X = 1, OP_V = 4, NOPS = 7 };
// Support points per instance (a 15x15 shell and its window fit). The
// per-support arrays of SolveDesc are narrow integers so that the descriptor
// stays well inside the 4 KiB of kernel arguments.
constexpr int MAX_SUPPORT = 256;
constexpr int STEP_HEAD = 12;     // doubles before the phases in a steptab row

// Role of a support point in the Hermitian-paired sums: when every block
// satisfies B(-n) = B(n)^dagger, sum_s w_s B_s = B_0 + sum_reps [w B + conj(w) B^dagger]
// and only the origin and one representative of each +-n pair are visited.
enum { ROLE_ORIGIN = 0, ROLE_REP = 1, ROLE_PARTNER = 2 };

// Row-separable sums: with k = m1 N2 + m2 and R_s = n1_s a1 + n2_s a2,
//   sum_s B_s e^{-i(k-A).R_s} = sum_d tw2[m2 n2_d] G_d(m1, t),
//   G_d = sum_{s: n2_s = n2_d} tw1[m1 n1_s] e^{iA.R_s} B_s,
// so the supports are grouped by their distinct n2 values.
constexpr int SEP_MAXD = 8;

struct SolveDesc {  // a kernel argument: at most 4 KiB with PairDesc and InstPtrs
    int nb, N1, N2, N, nt, ns;
    int stride;                   // doubles per steptab row = STEP_HEAD + 2*ns
    int xmask, vmask;             // which rbc / v axes are live
    int herm;                     // 1 when the paired (Hermitian) sums are exact
    int16_t n1[MAX_SUPPORT], n2[MAX_SUPPORT];
    uint8_t opmask[MAX_SUPPORT];  // bit o set when block (s, o) is non-zero (NOPS = 7 bits)
    uint8_t role[MAX_SUPPORT];    // ROLE_* (meaningful when herm == 1)
    int hermhx;                   // H and rbc blocks satisfy B(-n) = B(n)^dagger
    int nd;                       // distinct n2 values of the supports (0 when > SEP_MAXD)
    int dn2[SEP_MAXD];            // those values
    int dseg[SEP_MAXD + 1];       // supports of value d: sorder[dseg[d] .. dseg[d + 1])
    uint16_t sorder[MAX_SUPPORT];
    int nwd;                      // values d with a velocity block; wdi[d] = their rank or -1
    int wdi[SEP_MAXD];
};

// Pair pack of the Hermitian nb = 2 kernel. A +-R pair of blocks B, B^dagger
// with phases w, conj(w) adds to the Hermitian 2x2 sum [h00 h01; h01* h11]
//   h00 += 2 Re(w B00), h11 += 2 Re(w B11), h01 += w B01 + conj(w B10),
// i.e. four real outputs that are linear in (Re w, Im w): 8 coefficients per
// operator and pair. coef layout (doubles):
//   [0, 4)                 origin H: (B00.re, B11.re, B01.re, B01.im)
//   [4 + 4x, 8 + 4x)       origin rbc, slot x < 3
//   [16 + 4v, 20 + 4v)     origin v, slot v < 3, the 01 entries doubled
//   PAIR_HEAD + PAIR_REP r: per pair r: H (8), rbc slots (3 x 8), v slots (3 x 8)
// Coefficient c = 2*j + {0: Re w, 1: Im w} of output j. The v outputs are
// pre-scaled so that J = m0 p + m1 q + m2 Re(c) + m3 Im(c) with rho =
// [p c; c* q]. steps rows: 9 stage fields (3 stages x 3 rbc slots), 1 pad,
// then e^{i A(t_i).R_r} per pair r (complex).
constexpr int PAIR_MAX = 32;
constexpr int PAIR_HEAD = 28;
constexpr int PAIR_REP = 56;
constexpr int PAIR_STEP_HEAD = 10;

// Streamed pack (k_pair2s): H and rbc pairs as above; the current may be
// complex (velocity blocks whose -R partner lies outside the 4x4 window).
// With rho_vec = (p, q, Re c, Im c), J = sum_j rho_vec_j (m_j + i m'_j) and
// m_j = o_j + sum_r (a_rj Re w + b_rj Im w), m'_j likewise. Layout (doubles):
//   [0, 4) H0, [4, 16) rbc0 slots, [16, 40) v0 slots: (o_0..3, o'_0..3)
//   PS_HEAD + PS_REP r: H (8), rbc slots (3 x 8), v slots (3 x 16: a0 b0 .. a3 b3, a'0 b'0 .. a'3 b'3)
constexpr int PS_HEAD = 40;
constexpr int PS_REP = 80;

// Streamed Bloch-vector pack (k_bloch2s: 2 bands, the same occupation at every
// k), derived from the streamed pack at upload. With rho = r0 I + s.sigma,
// s = (Re c, -Im c, (p - q) / 2), a Hermitian output set (a, d, Re b, Im b)
// maps to g = (2 Re b, -2 Im b, a - d) and a current (m_0..3) to
// u = (m_2, -m_3, m_0 - m_1), J = u.s + r0 (m_0 + m_1). Layout (doubles):
//   [0, 3) g0, [3, 12) rbc0 slots, [12, 21) u0 slots (Re J), [21, 30) u0' (Im J)
//   BS_HEAD + BS_REP r: g (3 x 2), rbc slots (3 x 3 x 2), v slots (3 x (Re J 3 x 2, Im J 3 x 2))
// Coefficient 2*j + {0: Re w, 1: Im w} of component j.
constexpr int BS_HEAD = 32;
constexpr int BS_REP = 60;
constexpr int BS_N = BS_HEAD + PAIR_MAX * BS_REP;

struct PairDesc {
    int nrep, nx, nv, stride;
    int vpaired;          // 1: every v block has its -R partner, the current is real
    int xaxis[3], vaxis[3];
    int n1[PAIR_MAX], n2[PAIR_MAX];
};

// Device pointers and scalars of one instance (solve) on the GPU.
struct InstPtrs {
    const double *blk, *steptab, *tw1, *tw2, *pcoef, *scoef, *psteps, *rho_in;
    double *rho_out, *jpart;
    int nwaves;          // waves (current partials per step) of this instance's launch
    int kpw;             // k-points per wave of the row-separable kernel (16 bands: 1, 2 or 4)
    double dt;
};

// Descriptor of one instance of a batched launch, resident in HBM.
struct DevInst {
    SolveDesc d;
    PairDesc pd;
    InstPtrs p;
    double* J;           // [nt][6] reduced current
    int C;               // current components per wave partial
    int omap[6];         // output slot (re/im of x, y, z = 0..5) of each component, -1 = unused
    double jcst[6];      // constant added to each component after the 1/N normalisation
};


static_assert(sizeof(SolveDesc) + sizeof(PairDesc) + sizeof(InstPtrs) <= 4096 - 64,
              "descriptors must fit the kernel argument segment");
}  // namespace hhg
