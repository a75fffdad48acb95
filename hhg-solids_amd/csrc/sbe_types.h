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
constexpr int MAX_SUPPORT = 64;   // support points per instance (5x5 shell + window fits)
constexpr int STEP_HEAD = 12;     // doubles before the phases in a steptab row

struct SolveDesc {
    int nb, N1, N2, N, nt, ns;
    int stride;                   // doubles per steptab row = STEP_HEAD + 2*ns
    int xmask, vmask;             // which rbc / v axes are live
    int n1[MAX_SUPPORT], n2[MAX_SUPPORT];
    uint32_t opmask[MAX_SUPPORT]; // bit o set when block (s, o) is non-zero
};

}  // namespace hhg
