// sbe_kernels.hip -- CDNA4 (gfx950) kernels of the k-space SBE time loop.
//
// Every k-point is an independent nb x nb RK4 system (the per-k form of
// src/solver.cpp:86-181,1428-1726), so the whole time loop runs inside one
// launch: each k keeps its density matrix in registers for all nt steps and
// only the per-step current partials leave the CU. Per step and k:
//
//   w_s   = e^{-2 pi i (m1 n1_s/N1 + m2 n2_s/N2)} * e^{i A(t_i).R_s}
//   H~    = sum_s w_s H_s,   X~_a = sum_s w_s rbc_a,s      (Wannier Fourier sums)
//   K_j   = -i [H~ + E_j . X~, rho_j]                     (4 RK4 stages)
//   rho  += 0.1666666666666 dt (K1 + 2 K2 + 2 K3 + K4)
//   J_a  += sum_s w_s sum_ij v_a,s[ij] conj(rho_ij)       (4x4 window, src/observable.cpp:16-30)
//
// Kernels:
//   k_pair2<NREP,NX,NV,T>      nb = 2 with Hermitian blocks: one lane per k,
//                              the +-R pairs folded into real coefficient
//                              matrices, branch-free time loop.
//   k_pair2s<NREP,NX,NV,T>     nb = 2 with Hermitian H and rbc, pair
//                              coefficients streamed through scalar loads;
//                              complex (window) current.
//   k_sep<NBP,MAXD,T>          Hermitian 3 <= nb <= 16: row-separable Fourier
//                              sums built once per k-row and step in LDS,
//                              2x2 element blocks per lane, each lane group
//                              carrying 1, 2 or 4 k-points of the row per
//                              build, M - M^dagger commutator (16 bands: M as
//                              three real products on the matrix cores).
//   k_thread_per_k<NB,MAXU,T>  one lane per k, everything in registers; the
//                              blocks and the per-step table are uniform and
//                              arrive through scalar loads. For nb <= 2.
//   k_elem_per_lane<EPL,T>     G = pow2(nb*nb) <= 64 lanes per k (a whole
//                              workgroup beyond 16 bands), each lane owns EPL
//                              matrix elements; the stage Hamiltonian and
//                              stage rho are staged through LDS for the
//                              commutator. Any model, nb <= 32.
//   k_reduce_j                 deterministic sum of the per-wave partials.
//   k_init_state_lds           initial rho (H_k, Jacobi, T = 0 projector) for
//                              3 <= nb <= 32, a 16- or 32-lane group per k in
//                              LDS; k_init_state (thread per k) for nb <= 2.
// T is the stage arithmetic (double for the plain path, float for the mixed
// path); the state rho and the current are always FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "sbe_host.h"

namespace hhg {

#define HIP_CHECK(x)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess)                                                               \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) +   \
                                     " at " __FILE__ ":" + std::to_string(__LINE__));       \
    } while (0)

// Read-only tables with wave-uniform addresses are read through the constant
// address space, which lets the compiler use scalar (SMEM) loads.
typedef const __attribute__((address_space(4))) double* cdptr;
typedef const __attribute__((address_space(4))) float* cfptr;
__device__ __forceinline__ cdptr as_const(const double* p) { return (cdptr)(p); }
// Streamed pair coefficients: SC_N doubles, then the same table as floats
// (read by the FP32 kernels, which would otherwise convert every value).
constexpr int SC_N = PS_HEAD + PAIR_MAX * PS_REP;
// The streamed Bloch-vector pack follows in the same buffer: BS_N doubles at
// BS_OFF, then the same table as floats.
constexpr int BS_OFF = SC_N + (SC_N + 1) / 2;
// An opaque zero: offsets a per-step load so it is not hoisted out of the loop.
__device__ __forceinline__ int opaque_zero() {
    int z = 0;
    asm volatile("" : "+s"(z));
    return z;
}

template <typename T>
struct cx {
    T re, im;
};
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ void cmac(cx<T>& acc, cx<T> a, cx<T> b) {  // acc += a*b
    acc.re += a.re * b.re - a.im * b.im;
    acc.im += a.re * b.im + a.im * b.re;
}
template <typename T>
__device__ __forceinline__ void cmsub(cx<T>& acc, cx<T> a, cx<T> b) {  // acc -= a*b
    acc.re -= a.re * b.re - a.im * b.im;
    acc.im -= a.re * b.im + a.im * b.re;
}
template <typename T, typename S>
__device__ __forceinline__ cx<T> cast(cx<S> a) {
    return {(T)a.re, (T)a.im};
}
template <typename T>
__device__ __forceinline__ cx<T> ld(const double* p) {  // complex from interleaved doubles
    return {(T)p[0], (T)p[1]};
}
template <typename T>
__device__ __forceinline__ cx<T> ld(cdptr p) {
    return {(T)p[0], (T)p[1]};
}

__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int posmod(long a, int n) {
    int r = (int)(a % n);
    return r < 0 ? r + n : r;
}

// Deferred per-wave reduction of the current. Each lane parks its C per-step
// components for TCH steps in an LDS row (padded to avoid bank conflicts);
// every TCH steps lane p sums component p of the chunk over the 64 rows in a
// fixed order and writes the wave partial. Replaces a shuffle tree per step.
template <int C, int TCH>
struct JStage {
    static_assert(C * TCH <= 64, "one lane per (step, component) of a chunk");
    static constexpr int ROW = C * TCH + 1;
    static constexpr size_t BYTES = sizeof(double) * 64 * ROW;
    __device__ static void put(double* lds, int lane, int slot, const double* v) {
#pragma unroll
        for (int c = 0; c < C; ++c) lds[lane * ROW + slot * C + c] = v[c];
    }
    // steps t0 .. t0+nsteps-1 are parked; jpart is [nt][nwaves][C].
    // WG: the workgroup is this one wave and synchronises with barriers; else
    // the LDS row block belongs to one wave of a larger workgroup, which
    // orders its own LDS accesses (in issue order per wave) with wave fences.
    template <bool WG = true>
    __device__ static void flush(const double* lds, int lane, int t0, int nsteps, double* jpart, int nwaves,
                                 int wave) {
        auto sync = [] {
            if constexpr (WG) {
                __syncthreads();
            } else {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        };
        sync();
        if (lane < C * nsteps) {
            double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll 4
            for (int l = 0; l < 64; l += 4) {
                a0 += lds[l * ROW + lane];
                a1 += lds[(l + 1) * ROW + lane];
                a2 += lds[(l + 2) * ROW + lane];
                a3 += lds[(l + 3) * ROW + lane];
            }
            jpart[((size_t)(t0 + lane / C) * nwaves + wave) * C + lane % C] = (a0 + a1) + (a2 + a3);
        }
        sync();
    }
};

// --------------------------------------------------------------------------
// One lane per k. MAXU bounds the support size so that the per-k DFT phases
// live in registers.
template <int NB, int MAXU, typename T>
__device__ __forceinline__ void thread_per_k_body(const SolveDesc& d, const InstPtrs& ip, int bid) {
    const double* __restrict__ blk = ip.blk;
    const double* __restrict__ steptab = ip.steptab;
    const double* __restrict__ tw1 = ip.tw1;
    const double* __restrict__ tw2 = ip.tw2;
    const double* __restrict__ rho_in = ip.rho_in;
    double* __restrict__ rho_out = ip.rho_out;
    double* __restrict__ jpart = ip.jpart;
    const int nwaves = ip.nwaves;
    const double dt = ip.dt;
    constexpr int E = NB * NB;
    const int k = bid * 64 + threadIdx.x;
    const bool active = k < d.N;
    const int kk = active ? k : 0;
    const int m1 = kk / d.N2, m2 = kk - (kk / d.N2) * d.N2;
    const int wave = bid;
    const int lane = threadIdx.x & 63;

    cx<double> w0[MAXU];
#pragma unroll
    for (int s = 0; s < MAXU; ++s) {
        if (s < d.ns) {
            const int j1 = posmod((long)m1 * d.n1[s], d.N1), j2 = posmod((long)m2 * d.n2[s], d.N2);
            w0[s] = cmul(ld<double>(tw1 + 2 * j1), ld<double>(tw2 + 2 * j2));
        } else {
            w0[s] = {0.0, 0.0};
        }
    }
    cx<double> rho[E];
#pragma unroll
    for (int e = 0; e < E; ++e)
        rho[e] = active ? ld<double>(rho_in + 2 * ((size_t)e * d.N + k)) : cx<double>{0.0, 0.0};

    const T half = (T)(0.5 * dt);
    const T full = (T)dt;
    const double pref = 0.1666666666666 * dt;
    constexpr int TCH = 8;
    using JS = JStage<6, TCH>;
    __shared__ double jl[64 * JS::ROW];

    for (int ti = 0; ti < d.nt; ++ti) {
        const double* st = steptab + (size_t)ti * d.stride;
        cx<T> Ht[E], Xt[3][E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            Ht[e] = {0, 0};
            Xt[0][e] = Xt[1][e] = Xt[2][e] = {0, 0};
        }
#pragma unroll
        for (int s = 0; s < MAXU; ++s) {
            if (s < d.ns) {
                const uint32_t om = d.opmask[s];
                const cx<T> w = cast<T>(cmul(w0[s], ld<double>(st + STEP_HEAD + 2 * s)));
                const double* b = blk + (size_t)s * NOPS * E * 2;
                if (om & (1u << OP_H)) {
#pragma unroll
                    for (int e = 0; e < E; ++e) cmac(Ht[e], w, ld<T>(b + 2 * (OP_H * E + e)));
                }
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    if (om & (1u << (OP_X + a))) {
#pragma unroll
                        for (int e = 0; e < E; ++e) cmac(Xt[a][e], w, ld<T>(b + 2 * ((OP_X + a) * E + e)));
                    }
            }
        }
        cx<T> Hs[E], rs[E], K[E], acc[E];
        auto stage_h = [&](int j) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                cx<T> h = Ht[e];
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    if (d.xmask & (1 << a)) {
                        const T f = (T)st[3 * j + a];
                        h.re += f * Xt[a][e].re;
                        h.im += f * Xt[a][e].im;
                    }
                Hs[e] = h;
            }
        };
        auto deriv = [&]() {  // K = -i (Hs rs - rs Hs)
#pragma unroll
            for (int i = 0; i < NB; ++i)
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    cx<T> c = {0, 0};
#pragma unroll
                    for (int l = 0; l < NB; ++l) {
                        cmac(c, Hs[i * NB + l], rs[l * NB + j]);
                        cmsub(c, rs[i * NB + l], Hs[l * NB + j]);
                    }
                    K[i * NB + j] = {c.im, -c.re};
                }
        };
        stage_h(0);
#pragma unroll
        for (int e = 0; e < E; ++e) rs[e] = cast<T>(rho[e]);
        deriv();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            acc[e] = K[e];
            rs[e] = {(T)rho[e].re + half * K[e].re, (T)rho[e].im + half * K[e].im};
        }
        stage_h(1);
        deriv();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            acc[e].re += (T)2 * K[e].re; acc[e].im += (T)2 * K[e].im;
            rs[e] = {(T)rho[e].re + half * K[e].re, (T)rho[e].im + half * K[e].im};
        }
        deriv();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            acc[e].re += (T)2 * K[e].re; acc[e].im += (T)2 * K[e].im;
            rs[e] = {(T)rho[e].re + full * K[e].re, (T)rho[e].im + full * K[e].im};
        }
        stage_h(2);
        deriv();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            acc[e].re += K[e].re; acc[e].im += K[e].im;
            rho[e].re += pref * (double)acc[e].re;
            rho[e].im += pref * (double)acc[e].im;
        }
        // current of the new state with the phases of step ti
        double jc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < MAXU; ++s) {
            if (s < d.ns) {
                const uint32_t om = d.opmask[s];
                if (!(om & (7u << OP_V))) continue;
                const cx<double> w = cmul(w0[s], ld<double>(st + STEP_HEAD + 2 * s));
                const double* b = blk + (size_t)s * NOPS * E * 2;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    if (om & (1u << (OP_V + a))) {
                        cx<double> part = {0.0, 0.0};
#pragma unroll
                        for (int e = 0; e < E; ++e)
                            cmac(part, ld<double>(b + 2 * ((OP_V + a) * E + e)), cx<double>{rho[e].re, -rho[e].im});
                        const cx<double> c = cmul(part, w);
                        jc[2 * a] += c.re;
                        jc[2 * a + 1] += c.im;
                    }
            }
        }
        JS::put(jl, lane, ti % TCH, jc);
        if (ti % TCH == TCH - 1 || ti == d.nt - 1) JS::flush(jl, lane, ti - ti % TCH, ti % TCH + 1, jpart, nwaves, wave);
    }
    if (active) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            rho_out[2 * ((size_t)e * d.N + k)] = rho[e].re;
            rho_out[2 * ((size_t)e * d.N + k) + 1] = rho[e].im;
        }
    }
}

// --------------------------------------------------------------------------
// nb = 2 with Hermitian blocks (B(-R) = B(R)^dagger for every operator): one
// lane per k, rho kept as (p = rho00, q = rho11, c = rho01), rho10 = conj(c).
// Each +-R pair enters through its 4x2 real coefficient matrix (PairDesc):
// 8 FMAs per operator and pair and step, no branches. The commutator of two
// Hermitian 2x2 matrices H = [a b; b* d] and rho is
//   K00 = -K11 = 2 Im(b c*),   K01 = -i ((a - d) c + (q - p) b),
// and the current is real: J = m0 p + m1 q + m2 Re(c) + m3 Im(c).
// The coefficients of the NREP pairs, with this k's twiddle e^{-i k.R} folded
// in, stay in registers for the whole time loop; a step reads only the uniform
// pair phases and field components of its row of the step table.
template <int NREP, int NX, int NV, typename T>
__device__ __forceinline__ void pair2_body(const SolveDesc& d, const PairDesc& pd, const InstPtrs& ip, int bid) {
    const cdptr coef = as_const(ip.pcoef);
    const double* __restrict__ tw1 = ip.tw1;
    const double* __restrict__ tw2 = ip.tw2;
    const double* __restrict__ rho_in = ip.rho_in;
    double* __restrict__ rho_out = ip.rho_out;
    double* __restrict__ jpart = ip.jpart;
    const int nwaves = ip.nwaves;
    const double dt = ip.dt;
    const int k = bid * 64 + threadIdx.x;
    const bool active = k < d.N;
    const int kk = active ? k : 0;
    const int m1 = kk / d.N2, m2 = kk - (kk / d.N2) * d.N2;
    const int wave = bid;
    const int lane = threadIdx.x & 63;
    constexpr int TCH = 16;
    using JS = JStage<3, TCH>;
    __shared__ double jl[64 * JS::ROW];
    constexpr int NXA = NX > 0 ? NX : 1;

    // Coefficients in VGPRs for the whole time loop, with this k's twiddle
    // w0 = e^{-i k.R} folded in: a pair term c0 Re(w) + c1 Im(w), w = w0 phi(t),
    // equals phi_r (c0 w0r + c1 w0i) + phi_i (c1 w0r - c0 w0i), so a step needs
    // the uniform phase phi of each pair and no per-k complex product.
    const double lz = 0.0 * (double)lane;  // lane-varying zero for the uniform origin terms
    T hc[NREP][8], xc[NREP][NXA][8];
    double vc[NREP][NV][8];
#pragma unroll
    for (int r = 0; r < NREP; ++r) {
        const int j1 = posmod((long)m1 * pd.n1[r], d.N1), j2 = posmod((long)m2 * pd.n2[r], d.N2);
        const cx<double> w0 = cmul(ld<double>(tw1 + 2 * j1), ld<double>(tw2 + 2 * j2));
        const cdptr cr = coef + PAIR_HEAD + PAIR_REP * r;
        auto fold = [&](const double c0, const double c1, double& a, double& b) {
            a = c0 * w0.re + c1 * w0.im;
            b = c1 * w0.re - c0 * w0.im;
        };
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double a, b;
            fold(cr[2 * j], cr[2 * j + 1], a, b);
            hc[r][2 * j] = (T)a; hc[r][2 * j + 1] = (T)b;
#pragma unroll
            for (int x = 0; x < NX; ++x) {
                fold(cr[8 + 8 * x + 2 * j], cr[8 + 8 * x + 2 * j + 1], a, b);
                xc[r][x][2 * j] = (T)a; xc[r][x][2 * j + 1] = (T)b;
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) fold(cr[32 + 8 * v + 2 * j], cr[32 + 8 * v + 2 * j + 1], vc[r][v][2 * j], vc[r][v][2 * j + 1]);
        }
    }
    T h0[4], x0[NXA][4];
    double v0[NV][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        h0[j] = (T)(coef[j] + lz);
#pragma unroll
        for (int x = 0; x < NX; ++x) x0[x][j] = (T)(coef[4 + 4 * x + j] + lz);
#pragma unroll
        for (int v = 0; v < NV; ++v) v0[v][j] = coef[16 + 4 * v + j] + lz;
    }

    double p = 0, q = 0;
    cx<double> c = {0, 0};
    if (active) {
        p = rho_in[0 * 2 * (size_t)d.N + 2 * k];
        c = ld<double>(rho_in + 1 * 2 * (size_t)d.N + 2 * k);
        q = rho_in[3 * 2 * (size_t)d.N + 2 * k];
    }
    const T half = (T)(0.5 * dt);
    const T full = (T)dt;
    const double pref = 0.1666666666666 * dt;

    // State of a step: its row of the step table (pair phases phi, field) and
    // the field-free Fourier sums of H and rbc. Two buffers alternate: step ti
    // reads its own and fills the other one for step ti + 1 -- the row load is
    // issued at the top of the step and the sums are formed at its end, so the
    // load has the whole step to arrive. The time loop is unrolled by two so
    // that the alternation needs no register copies.
    struct Buf {
        double ph[NREP][2], ef[3][NXA];
        T hs[4], xs[NXA][4];
    };
    auto sums = [&](Buf& b) {
#pragma unroll
        for (int j = 0; j < 4; ++j) b.hs[j] = h0[j];
#pragma unroll
        for (int r = 0; r < NREP; ++r) {
            const T wr = (T)b.ph[r][0], wi = (T)b.ph[r][1];
#pragma unroll
            for (int j = 0; j < 4; ++j) b.hs[j] = fmaT(hc[r][2 * j + 1], wi, fmaT(hc[r][2 * j], wr, b.hs[j]));
        }
#pragma unroll
        for (int x = 0; x < NX; ++x) {
#pragma unroll
            for (int j = 0; j < 4; ++j) b.xs[x][j] = x0[x][j];
#pragma unroll
            for (int r = 0; r < NREP; ++r) {
                const T wr = (T)b.ph[r][0], wi = (T)b.ph[r][1];
#pragma unroll
                for (int j = 0; j < 4; ++j) b.xs[x][j] = fmaT(xc[r][x][2 * j + 1], wi, fmaT(xc[r][x][2 * j], wr, b.xs[x][j]));
            }
        }
    };
    // Rows arrive as vector loads (lane-varying zero offset; nt is never
    // negative): scalar loads share their wait counter with the LDS traffic of
    // the current staging and would be drained with it.
    const int zl = lane * (int)(d.nt < 0);
    auto load_row = [&](int ti, Buf& b) {
        const double* row = ip.psteps + (size_t)ti * pd.stride + zl;
#pragma unroll
        for (int r = 0; r < NREP; ++r) { b.ph[r][0] = row[PAIR_STEP_HEAD + 2 * r]; b.ph[r][1] = row[PAIR_STEP_HEAD + 2 * r + 1]; }
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int x = 0; x < NXA; ++x) b.ef[j][x] = NX > 0 ? row[3 * j + x] : 0.0;
    };

    // chk: the step reduces the previous chunk of TCH steps when it starts a
    // chunk (at its top, so that the stage chain, the current and the next
    // sums form one block)
    auto step = [&](int ti, const Buf& cur, Buf& nxt, bool chk = true) {
        if (chk && ti % TCH == 0 && ti > 0) JS::flush(jl, lane, ti - TCH, TCH, jpart, nwaves, wave);
        load_row(ti + 1 < d.nt ? ti + 1 : ti, nxt);
        T A, D, Br, Bi;
        auto stage_h = [&](int j) {
            A = cur.hs[0]; D = cur.hs[1]; Br = cur.hs[2]; Bi = cur.hs[3];
#pragma unroll
            for (int x = 0; x < NX; ++x) {
                const T f = (T)cur.ef[j][x];
                A = fmaT(f, cur.xs[x][0], A); D = fmaT(f, cur.xs[x][1], D);
                Br = fmaT(f, cur.xs[x][2], Br); Bi = fmaT(f, cur.xs[x][3], Bi);
            }
        };
        T kp, kr, ki;
        auto deriv = [&](T sp, T sq, T scr, T sci) {
            kp = (T)2 * fmaT(Bi, scr, -(Br * sci));
            const T ad = A - D, qp = sq - sp;
            kr = fmaT(ad, sci, qp * Bi);
            ki = -fmaT(ad, scr, qp * Br);
        };
        const T P = (T)p, Q = (T)q, Cr = (T)c.re, Ci = (T)c.im;
        stage_h(0);
        deriv(P, Q, Cr, Ci);
        T ap = kp, ar = kr, ai = ki;
        stage_h(1);
        deriv(fmaT(half, kp, P), fmaT(-half, kp, Q), fmaT(half, kr, Cr), fmaT(half, ki, Ci));
        ap = fmaT((T)2, kp, ap); ar = fmaT((T)2, kr, ar); ai = fmaT((T)2, ki, ai);
        deriv(fmaT(half, kp, P), fmaT(-half, kp, Q), fmaT(half, kr, Cr), fmaT(half, ki, Ci));
        ap = fmaT((T)2, kp, ap); ar = fmaT((T)2, kr, ar); ai = fmaT((T)2, ki, ai);
        stage_h(2);
        deriv(fmaT(full, kp, P), fmaT(-full, kp, Q), fmaT(full, kr, Cr), fmaT(full, ki, Ci));
        ap += kp; ar += kr; ai += ki;
        p = fmaT(pref, (double)ap, p);
        q = fmaT(-pref, (double)ap, q);
        c.re = fmaT(pref, (double)ar, c.re);
        c.im = fmaT(pref, (double)ai, c.im);
        // current of the new state with the phases of step ti, FP64
        double jc[3] = {0, 0, 0};
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double m[4] = {v0[v][0], v0[v][1], v0[v][2], v0[v][3]};
#pragma unroll
            for (int r = 0; r < NREP; ++r)
#pragma unroll
                for (int j = 0; j < 4; ++j) m[j] = fmaT(vc[r][v][2 * j + 1], cur.ph[r][1], fmaT(vc[r][v][2 * j], cur.ph[r][0], m[j]));
            // inactive lanes hold rho = 0 and contribute exactly 0
            jc[v] = fmaT(m[3], c.im, fmaT(m[2], c.re, fmaT(m[1], q, m[0] * p)));
        }
        JS::put(jl, lane, ti % TCH, jc);
        sums(nxt);
    };

    Buf B0, B1;
    load_row(0, B0);
    sums(B0);
    // VGPRs held by the per-k coefficients; the unrolled loop keeps two step
    // states live and is used while the total stays within the 256 VGPRs
    constexpr int CREGS = NREP * 8 * (1 + NX) * (int)(sizeof(T) / 4) + NREP * NV * 16;
    if constexpr (CREGS <= 160) {
        int ti = 0;
        // whole chunks of TCH steps with no branch inside, as in bloch2_body
        for (; ti + TCH <= d.nt; ti += TCH) {
            if (ti > 0) JS::flush(jl, lane, ti - TCH, TCH, jpart, nwaves, wave);
#pragma unroll
            for (int u = 0; u < TCH; u += 2) {
                step(ti + u, B0, B1, false);
                step(ti + u + 1, B1, B0, false);
            }
        }
        for (; ti + 2 <= d.nt; ti += 2) {
            step(ti, B0, B1);
            step(ti + 1, B1, B0);
        }
        if (ti < d.nt) step(ti, B0, B1);
    } else {
        for (int ti = 0; ti < d.nt; ++ti) {
            step(ti, B0, B1);
            B0 = B1;
        }
    }
    if (d.nt > 0) {
        const int t0 = ((d.nt - 1) / TCH) * TCH;  // the last, possibly partial chunk
        JS::flush(jl, lane, t0, d.nt - t0, jpart, nwaves, wave);
    }
    if (active) {
        double* o = rho_out + 2 * (size_t)k;
        const size_t st = 2 * (size_t)d.N;
        o[0] = p; o[1] = 0.0;
        o[st] = c.re; o[st + 1] = c.im;
        o[2 * st] = c.re; o[2 * st + 1] = -c.im;
        o[3 * st] = q; o[3 * st + 1] = 0.0;
    }
}

// --------------------------------------------------------------------------
// nb = 2 with Hermitian blocks and the same trace at every k (a filled band
// at T = 0): the Bloch-vector form of the register-pair kernel. With
// rho = r0 I + s.sigma and H = h0 I + h.sigma, -i [H, rho] = (2 h x s).sigma,
// so r0 is constant, h0 drops out of the dynamics and a stage is one cross
// product g x s with g = 2 h = (2 Re b, -2 Im b, a - d) for H = [a b; b* d]
// (rho = [p c; c* q]: s = (Re c, -Im c, (p - q)/2), r0 = (p + q)/2). The
// current of a Hermitian window, J = m0 p + m1 q + m2 Re c + m3 Im c, splits
// into r0 (m0 + m1) -- summed over the grid this is the constant
// N r0 (m0 + m1) of the origin, since the k-sum of e^{-i k.R} vanishes for
// R != 0, added by the reduction -- and (m0 - m1) s_z + m2 s_x - m3 s_y.
// Per k-point and step: 3 outputs per Fourier sum instead of 4, 6 VALU
// operations per stage derivative instead of ~9, 3 state components.
// Per pair and operator the 4 outputs (a, d, Re b, Im b) have coefficients
// (c0, c1) of (Re w, Im w); with the twiddle w0 = e^{-i k.R} of the lane's
// k-point folded in (w = w0 phi) and mapped to the Bloch components:
//   g = (2 Re b, -2 Im b, a - d),   current u = (m2, -m3, m0 - m1).
// g0, x0, u0: the origin terms (coefficients of the constant phase 1).
template <int NREP, int NX, int NV, typename T>
struct B2Coef {
    static constexpr int NXA = NX > 0 ? NX : 1;
    T hc[NREP][6], xc[NREP][NXA][6];
    double vc[NREP][NV][6];
    T g0[3], x0[NXA][3];
    double u0[NV][3];
    __device__ __forceinline__ void init(const SolveDesc& d, const PairDesc& pd, const InstPtrs& ip, int m1, int m2,
                                         int lane) {
        const cdptr coef = as_const(ip.pcoef);
        const double lz = 0.0 * (double)lane;
        auto fold = [](const double c0, const double c1, const cx<double> w0, double& a, double& b) {
            a = c0 * w0.re + c1 * w0.im;
            b = c1 * w0.re - c0 * w0.im;
        };
        // 4 outputs x 2 folded coefficients -> 3 Bloch components x 2
        auto bloch = [&](const cdptr c, const cx<double> w0, const double sx, const double sy, double* out) {
            double f[4][2];
#pragma unroll
            for (int j = 0; j < 4; ++j) fold(c[2 * j], c[2 * j + 1], w0, f[j][0], f[j][1]);
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                out[u] = sx * f[2][u];
                out[2 + u] = sy * f[3][u];
                out[4 + u] = f[0][u] - f[1][u];
            }
        };
#pragma unroll
        for (int r = 0; r < NREP; ++r) {
            const int j1 = posmod((long)m1 * pd.n1[r], d.N1), j2 = posmod((long)m2 * pd.n2[r], d.N2);
            const cx<double> w0 = cmul(ld<double>(ip.tw1 + 2 * j1), ld<double>(ip.tw2 + 2 * j2));
            const cdptr cr = coef + PAIR_HEAD + PAIR_REP * r;
            double o[6];
            bloch(cr, w0, 2.0, -2.0, o);
#pragma unroll
            for (int u = 0; u < 6; ++u) hc[r][u] = (T)o[u];
#pragma unroll
            for (int x = 0; x < NX; ++x) {
                bloch(cr + 8 + 8 * x, w0, 2.0, -2.0, o);
#pragma unroll
                for (int u = 0; u < 6; ++u) xc[r][x][u] = (T)o[u];
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) bloch(cr + 32 + 8 * v, w0, 1.0, -1.0, vc[r][v]);
        }
        g0[0] = (T)(2.0 * coef[2] + lz); g0[1] = (T)(-2.0 * coef[3] + lz); g0[2] = (T)(coef[0] - coef[1] + lz);
#pragma unroll
        for (int x = 0; x < NX; ++x) {
            const cdptr c = coef + 4 + 4 * x;
            x0[x][0] = (T)(2.0 * c[2] + lz); x0[x][1] = (T)(-2.0 * c[3] + lz); x0[x][2] = (T)(c[0] - c[1] + lz);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const cdptr c = coef + 16 + 4 * v;
            u0[v][0] = c[2] + lz; u0[v][1] = -c[3] + lz; u0[v][2] = c[0] - c[1] + lz;
        }
    }
    // Fourier sums of one step row: g (gs) and the rbc vectors (xs)
    __device__ __forceinline__ void sums(const double (&ph)[NREP][2], T (&gs)[3], T (&xs)[NXA][3]) const {
#pragma unroll
        for (int j = 0; j < 3; ++j) gs[j] = g0[j];
#pragma unroll
        for (int r = 0; r < NREP; ++r) {
            const T wr = (T)ph[r][0], wi = (T)ph[r][1];
#pragma unroll
            for (int j = 0; j < 3; ++j) gs[j] = fmaT(hc[r][2 * j + 1], wi, fmaT(hc[r][2 * j], wr, gs[j]));
        }
#pragma unroll
        for (int x = 0; x < NX; ++x) {
#pragma unroll
            for (int j = 0; j < 3; ++j) xs[x][j] = x0[x][j];
#pragma unroll
            for (int r = 0; r < NREP; ++r) {
                const T wr = (T)ph[r][0], wi = (T)ph[r][1];
#pragma unroll
                for (int j = 0; j < 3; ++j) xs[x][j] = fmaT(xc[r][x][2 * j + 1], wi, fmaT(xc[r][x][2 * j], wr, xs[x][j]));
            }
        }
    }
    // current of state s with the phases of one step (FP64), without the r0 term
    __device__ __forceinline__ void current(const double (&ph)[NREP][2], double sx, double sy, double sz,
                                            double (&jc)[3]) const {
        jc[0] = jc[1] = jc[2] = 0.0;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double u[3] = {u0[v][0], u0[v][1], u0[v][2]};
#pragma unroll
            for (int r = 0; r < NREP; ++r)
#pragma unroll
                for (int j = 0; j < 3; ++j) u[j] = fmaT(vc[r][v][2 * j + 1], ph[r][1], fmaT(vc[r][v][2 * j], ph[r][0], u[j]));
            jc[v] = fmaT(u[2], sz, fmaT(u[1], sy, u[0] * sx));
        }
    }
};

template <int NREP, int NX, int NV, typename T>
__device__ __forceinline__ void bloch2_body(const SolveDesc& d, const PairDesc& pd, const InstPtrs& ip, int bid) {
    const double* __restrict__ rho_in = ip.rho_in;
    double* __restrict__ rho_out = ip.rho_out;
    double* __restrict__ jpart = ip.jpart;
    const int nwaves = ip.nwaves;
    const double dt = ip.dt;
    const int k = bid * 64 + threadIdx.x;
    const bool active = k < d.N;
    const int kk = active ? k : 0;
    const int m1 = kk / d.N2, m2 = kk - (kk / d.N2) * d.N2;
    const int wave = bid;
    const int lane = threadIdx.x & 63;
    constexpr int TCH = 16;
    using JS = JStage<3, TCH>;
    __shared__ double jl[64 * JS::ROW];
    constexpr int NXA = NX > 0 ? NX : 1;

    B2Coef<NREP, NX, NV, T> cf;
    cf.init(d, pd, ip, m1, m2, lane);

    double r0 = 0, sx = 0, sy = 0, sz = 0;
    if (active) {
        const double p = rho_in[2 * (size_t)k];
        const cx<double> c = ld<double>(rho_in + 1 * 2 * (size_t)d.N + 2 * k);
        const double q = rho_in[3 * 2 * (size_t)d.N + 2 * k];
        r0 = 0.5 * (p + q);
        sz = 0.5 * (p - q);
        sx = c.re;
        sy = -c.im;
    }
    const T half = (T)(0.5 * dt);
    const T full = (T)dt;
    const double pref = 0.1666666666666 * dt;

    // Step state: the row of the step table (pair phases, field) and the
    // field-free Fourier sums g and X; two buffers alternate as in k_pair2.
    struct Buf {
        double ph[NREP][2], ef[3][NXA];
        T gs[3], xs[NXA][3];
    };
    auto sums = [&](Buf& b) { cf.sums(b.ph, b.gs, b.xs); };
    const int zl = lane * (int)(d.nt < 0);
    auto load_row = [&](int ti, Buf& b) {
        const double* row = ip.psteps + (size_t)ti * pd.stride + zl;
#pragma unroll
        for (int r = 0; r < NREP; ++r) { b.ph[r][0] = row[PAIR_STEP_HEAD + 2 * r]; b.ph[r][1] = row[PAIR_STEP_HEAD + 2 * r + 1]; }
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int x = 0; x < NXA; ++x) b.ef[j][x] = NX > 0 ? row[3 * j + x] : 0.0;
    };

    // chk: the step reduces the previous chunk of the current when it starts one
    auto step = [&](int ti, const Buf& cur, Buf& nxt, bool chk = true) {
        if (chk && ti % TCH == 0 && ti > 0) JS::flush(jl, lane, ti - TCH, TCH, jpart, nwaves, wave);
        load_row(ti + 1 < d.nt ? ti + 1 : ti, nxt);
        T gx, gy, gz;
        auto stage_g = [&](int j) {
            gx = cur.gs[0]; gy = cur.gs[1]; gz = cur.gs[2];
#pragma unroll
            for (int x = 0; x < NX; ++x) {
                const T f = (T)cur.ef[j][x];
                gx = fmaT(f, cur.xs[x][0], gx); gy = fmaT(f, cur.xs[x][1], gy); gz = fmaT(f, cur.xs[x][2], gz);
            }
        };
        T kx, ky, kz;
        auto deriv = [&](T ax, T ay, T az) {  // k = g x s
            kx = fmaT(gy, az, -(gz * ay));
            ky = fmaT(gz, ax, -(gx * az));
            kz = fmaT(gx, ay, -(gy * ax));
        };
        const T Sx = (T)sx, Sy = (T)sy, Sz = (T)sz;
        stage_g(0);
        deriv(Sx, Sy, Sz);
        T ax = kx, ay = ky, az = kz;
        stage_g(1);
        deriv(fmaT(half, kx, Sx), fmaT(half, ky, Sy), fmaT(half, kz, Sz));
        ax = fmaT((T)2, kx, ax); ay = fmaT((T)2, ky, ay); az = fmaT((T)2, kz, az);
        deriv(fmaT(half, kx, Sx), fmaT(half, ky, Sy), fmaT(half, kz, Sz));
        ax = fmaT((T)2, kx, ax); ay = fmaT((T)2, ky, ay); az = fmaT((T)2, kz, az);
        stage_g(2);
        deriv(fmaT(full, kx, Sx), fmaT(full, ky, Sy), fmaT(full, kz, Sz));
        ax += kx; ay += ky; az += kz;
        sx = fmaT(pref, (double)ax, sx);
        sy = fmaT(pref, (double)ay, sy);
        sz = fmaT(pref, (double)az, sz);
        // current of the new state with the phases of step ti (FP64), without
        // the r0 term; inactive lanes hold s = 0 and contribute exactly 0
        double jc[3];
        cf.current(cur.ph, sx, sy, sz, jc);
        JS::put(jl, lane, ti % TCH, jc);
        sums(nxt);
    };

    Buf B0, B1;
    load_row(0, B0);
    sums(B0);
    constexpr int CREGS = NREP * 6 * (1 + NX) * (int)(sizeof(T) / 4) + NREP * NV * 12;
    if constexpr (CREGS <= 160) {
        int ti = 0;
        // whole chunks of TCH steps with no branch inside: the current of one
        // step and the chain of the next are scheduled together
        for (; ti + TCH <= d.nt; ti += TCH) {
            if (ti > 0) JS::flush(jl, lane, ti - TCH, TCH, jpart, nwaves, wave);
#pragma unroll
            for (int u = 0; u < TCH; u += 2) {
                step(ti + u, B0, B1, false);
                step(ti + u + 1, B1, B0, false);
            }
        }
        for (; ti + 2 <= d.nt; ti += 2) {
            step(ti, B0, B1);
            step(ti + 1, B1, B0);
        }
        if (ti < d.nt) step(ti, B0, B1);
    } else {
        for (int ti = 0; ti < d.nt; ++ti) {
            step(ti, B0, B1);
            B0 = B1;
        }
    }
    if (d.nt > 0) {
        const int t0 = ((d.nt - 1) / TCH) * TCH;
        JS::flush(jl, lane, t0, d.nt - t0, jpart, nwaves, wave);
    }
    if (active) {
        double* o = rho_out + 2 * (size_t)k;
        const size_t st = 2 * (size_t)d.N;
        o[0] = r0 + sz; o[1] = 0.0;
        o[st] = sx; o[st + 1] = -sy;
        o[2 * st] = sx; o[2 * st + 1] = sy;
        o[3 * st] = r0 - sz; o[3 * st + 1] = 0.0;
    }
}

// Two-wave form of bloch2_body for grids of at most two groups of 64 k-points
// per CU, where the one-wave form leaves SIMDs idle and its lone wave on a
// SIMD exposes every latency of the step. Per group, wave 0 runs the step
// (Fourier sums of the next row, the RK4 chain) and leaves each new state in
// a ring of 2 SB LDS slots; wave 1 forms the current of those states one
// block of SB steps later. One barrier per block; each wave passes
// ceil(nt / SB) barriers. Every value is formed by the same operations in the
// same order as in bloch2_body, so state and current are bitwise the same.
template <int NREP, int NX, int NV, typename T>
__device__ __forceinline__ void bloch2w_body(const SolveDesc& d, const PairDesc& pd, const InstPtrs& ip, int bid) {
    constexpr int NXA = NX > 0 ? NX : 1;
    constexpr int TCH = 16;
    constexpr int SB = 8;  // steps per barrier
    using JS = JStage<3, TCH>;
    __shared__ double jl[64 * JS::ROW];
    __shared__ double sq[2 * SB][3][64];
    const int nt = d.nt;
    // Both waves run the steps in blocks of SB, unrolled (no branch inside a
    // block), each followed by the barrier, and the rest of the steps (fewer
    // than SB) followed by one more: ceil(nt / SB) barriers each. SB is even,
    // so the two alternating step states line up across blocks.
    auto run_steps = [&](auto&& stp) {
        int i = 0;
        for (; i + SB <= nt; i += SB) {
#pragma unroll
            for (int u = 0; u < SB; u += 2) {
                stp(i + u, 0);
                stp(i + u + 1, 1);
            }
            __syncthreads();
        }
        if (i < nt) {
            for (; i + 2 <= nt; i += 2) {
                stp(i, 0);
                stp(i + 1, 1);
            }
            if (i < nt) stp(i, 0);
            __syncthreads();
        }
    };
    const int lane = threadIdx.x & 63;
    const int k = bid * 64 + lane;
    const bool active = k < d.N;
    const int kk = active ? k : 0;
    const int m1 = kk / d.N2, m2 = kk - (kk / d.N2) * d.N2;
    // the role as a scalar: each wave takes exactly one branch, so both pass
    // the same barriers
    const int role = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    B2Coef<NREP, NX, NV, T> cf;
    cf.init(d, pd, ip, m1, m2, lane);
    // step rows are wave-uniform: scalar loads
    auto load_ph = [&](int ti, double (&ph)[NREP][2]) {
        const cdptr row = as_const(ip.psteps) + (size_t)ti * pd.stride;
#pragma unroll
        for (int r = 0; r < NREP; ++r) { ph[r][0] = row[PAIR_STEP_HEAD + 2 * r]; ph[r][1] = row[PAIR_STEP_HEAD + 2 * r + 1]; }
    };

    if (role == 1) {  // wave 1: the current
        // J(ti): the state after step ti (slot (ti + 1) % 2 SB) with the phases of row ti
        auto current = [&](int ti, const double (&ph)[NREP][2]) {
            const double* s = &sq[(ti + 1) % (2 * SB)][0][0];
            double jc[3];
            cf.current(ph, s[lane], s[64 + lane], s[128 + lane], jc);
            JS::put(jl, lane, ti % TCH, jc);
            if (ti % TCH == TCH - 1 && ti + 1 < nt)
                JS::template flush<false>(jl, lane, ti + 1 - TCH, TCH, ip.jpart, ip.nwaves, bid);
        };
        // step i: J(i - SB), whose state wave 0 wrote in the block before;
        // prefetches the phases of row i - SB + 1
        auto fstep = [&](int i, const double (&pcur)[NREP][2], double (&pnext)[NREP][2]) {
            if (i - SB + 1 >= 0) load_ph(i - SB + 1, pnext);
            if (i - SB >= 0) current(i - SB, pcur);
        };
        double pj[2][NREP][2] = {};
        run_steps([&](int i, int par) {
            if (par == 0) fstep(i, pj[0], pj[1]);
            else fstep(i, pj[1], pj[0]);
        });
        if (nt > 0) {
            for (int t = nt > SB ? nt - SB : 0; t < nt; ++t) {  // the states of the last block
                double ph[NREP][2];
                load_ph(t, ph);
                current(t, ph);
            }
            const int t0 = ((nt - 1) / TCH) * TCH;
            JS::template flush<false>(jl, lane, t0, nt - t0, ip.jpart, ip.nwaves, bid);
        }
        return;
    }

    // wave 0: the step
    double r0 = 0, sx = 0, sy = 0, sz = 0;
    if (active) {
        const double p = ip.rho_in[2 * (size_t)k];
        const cx<double> c = ld<double>(ip.rho_in + 1 * 2 * (size_t)d.N + 2 * k);
        const double q = ip.rho_in[3 * 2 * (size_t)d.N + 2 * k];
        r0 = 0.5 * (p + q);
        sz = 0.5 * (p - q);
        sx = c.re;
        sy = -c.im;
    }
    const T half = (T)(0.5 * ip.dt);
    const T full = (T)ip.dt;
    const double pref = 0.1666666666666 * ip.dt;
    struct Buf {
        double ph[NREP][2], ef[3][NXA];
        T gs[3], xs[NXA][3];
    };
    auto load_row = [&](int ti, Buf& b) {
        load_ph(ti, b.ph);
        const cdptr row = as_const(ip.psteps) + (size_t)ti * pd.stride;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int x = 0; x < NXA; ++x) b.ef[j][x] = NX > 0 ? row[3 * j + x] : 0.0;
    };
    auto cstep = [&](int ti, const Buf& cur, Buf& nxt) {
        load_row(ti + 1 < nt ? ti + 1 : ti, nxt);
        T gx, gy, gz;
        auto stage_g = [&](int j) {
            gx = cur.gs[0]; gy = cur.gs[1]; gz = cur.gs[2];
#pragma unroll
            for (int x = 0; x < NX; ++x) {
                const T f = (T)cur.ef[j][x];
                gx = fmaT(f, cur.xs[x][0], gx); gy = fmaT(f, cur.xs[x][1], gy); gz = fmaT(f, cur.xs[x][2], gz);
            }
        };
        T kx, ky, kz;
        auto deriv = [&](T ax, T ay, T az) {  // k = g x s
            kx = fmaT(gy, az, -(gz * ay));
            ky = fmaT(gz, ax, -(gx * az));
            kz = fmaT(gx, ay, -(gy * ax));
        };
        const T Sx = (T)sx, Sy = (T)sy, Sz = (T)sz;
        stage_g(0);
        deriv(Sx, Sy, Sz);
        T ax = kx, ay = ky, az = kz;
        stage_g(1);
        deriv(fmaT(half, kx, Sx), fmaT(half, ky, Sy), fmaT(half, kz, Sz));
        ax = fmaT((T)2, kx, ax); ay = fmaT((T)2, ky, ay); az = fmaT((T)2, kz, az);
        deriv(fmaT(half, kx, Sx), fmaT(half, ky, Sy), fmaT(half, kz, Sz));
        ax = fmaT((T)2, kx, ax); ay = fmaT((T)2, ky, ay); az = fmaT((T)2, kz, az);
        stage_g(2);
        deriv(fmaT(full, kx, Sx), fmaT(full, ky, Sy), fmaT(full, kz, Sz));
        ax += kx; ay += ky; az += kz;
        sx = fmaT(pref, (double)ax, sx);
        sy = fmaT(pref, (double)ay, sy);
        sz = fmaT(pref, (double)az, sz);
        double* s = &sq[(ti + 1) % (2 * SB)][0][0];
        s[lane] = sx;
        s[64 + lane] = sy;
        s[128 + lane] = sz;
        cf.sums(nxt.ph, nxt.gs, nxt.xs);
    };
    Buf B0, B1;
    load_row(0, B0);
    cf.sums(B0.ph, B0.gs, B0.xs);
    run_steps([&](int i, int par) {
        if (par == 0) cstep(i, B0, B1);
        else cstep(i, B1, B0);
    });
    if (active) {
        double* o = ip.rho_out + 2 * (size_t)k;
        const size_t st = 2 * (size_t)d.N;
        o[0] = r0 + sz; o[1] = 0.0;
        o[st] = sx; o[st + 1] = -sy;
        o[2 * st] = sx; o[2 * st + 1] = sy;
        o[3 * st] = r0 - sz; o[3 * st + 1] = 0.0;
    }
}

// --------------------------------------------------------------------------
// nb = 2, Hermitian H and rbc, current possibly complex: pairs streamed.
// The phases of up to NREP pairs stay in VGPRs; the coefficients (PS_REP per
// pair) arrive through scalar loads every step -- the coefficient pointer is
// laundered per step so that the loads are not hoisted out of the time loop
// into more SGPRs than exist. Unused pair slots have phase 0.
template <int NREP, int NX, int NV, typename T>
__device__ __forceinline__ void pair2s_body(const SolveDesc& d, const PairDesc& pd, const InstPtrs& ip, int bid) {
    const cdptr steps = as_const(ip.psteps);
    const double* __restrict__ tw1 = ip.tw1;
    const double* __restrict__ tw2 = ip.tw2;
    const double* __restrict__ rho_in = ip.rho_in;
    double* __restrict__ rho_out = ip.rho_out;
    double* __restrict__ jpart = ip.jpart;
    const int nwaves = ip.nwaves;
    const double dt = ip.dt;
    const int k = bid * 64 + threadIdx.x;
    const bool active = k < d.N;
    const int kk = active ? k : 0;
    const int m1 = kk / d.N2, m2 = kk - (kk / d.N2) * d.N2;
    const int wave = bid;
    const int lane = threadIdx.x & 63;
    constexpr int C = 2 * NV;
    constexpr int TCH = 64 / C;
    using JS = JStage<C, TCH>;
    __shared__ double jl[64 * JS::ROW];
    constexpr int NXA = NX > 0 ? NX : 1;

    cx<double> w0[NREP];
#pragma unroll
    for (int r = 0; r < NREP; ++r) {
        w0[r] = {0.0, 0.0};
        if (r < pd.nrep) {
            const int j1 = posmod((long)m1 * pd.n1[r], d.N1), j2 = posmod((long)m2 * pd.n2[r], d.N2);
            w0[r] = cmul(ld<double>(tw1 + 2 * j1), ld<double>(tw2 + 2 * j2));
        }
    }
    double p = 0, q = 0;
    cx<double> c = {0, 0};
    if (active) {
        p = rho_in[0 * 2 * (size_t)d.N + 2 * k];
        c = ld<double>(rho_in + 1 * 2 * (size_t)d.N + 2 * k);
        q = rho_in[3 * 2 * (size_t)d.N + 2 * k];
    }
    const T half = (T)(0.5 * dt);
    const T full = (T)dt;
    const double pref = 0.1666666666666 * dt;

    // chk: the step reduces its chunk of the current when it ends one
    auto step = [&](int ti, bool chk) {
        const cdptr row = steps + (size_t)ti * pd.stride;
        const cdptr cf = as_const(ip.scoef) + opaque_zero();
        const cfptr cf32 = (cfptr)(as_const(ip.scoef) + SC_N) + opaque_zero();
        auto cT = [&](int i) -> T {  // H and rbc coefficient i in the stage type
            if constexpr (std::is_same<T, float>::value) return cf32[i];
            else return (T)cf[i];
        };
        cx<double> w[NREP];
#pragma unroll
        for (int r = 0; r < NREP; ++r) w[r] = cmul(w0[r], ld<double>(row + PAIR_STEP_HEAD + 2 * r));
        T hs[4], xs[NXA][4];
        double mr[NV][4], mi[NV][4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            hs[j] = cT(j);
#pragma unroll
            for (int x = 0; x < NX; ++x) xs[x][j] = cT(4 + 4 * x + j);
#pragma unroll
            for (int v = 0; v < NV; ++v) { mr[v][j] = cf[16 + 8 * v + j]; mi[v][j] = cf[16 + 8 * v + 4 + j]; }
        }
#pragma unroll
        for (int r = 0; r < NREP; ++r) {
            const int o = PS_HEAD + PS_REP * r;
            const cdptr cr = cf + o;
            const T wr = (T)w[r].re, wi = (T)w[r].im;
#pragma unroll
            for (int j = 0; j < 4; ++j) hs[j] = fmaT(cT(o + 2 * j + 1), wi, fmaT(cT(o + 2 * j), wr, hs[j]));
#pragma unroll
            for (int x = 0; x < NX; ++x)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    xs[x][j] = fmaT(cT(o + 8 + 8 * x + 2 * j + 1), wi, fmaT(cT(o + 8 + 8 * x + 2 * j), wr, xs[x][j]));
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const cdptr cv = cr + 32 + 16 * v;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    mr[v][j] = fmaT(cv[2 * j + 1], w[r].im, fmaT(cv[2 * j], w[r].re, mr[v][j]));
                    mi[v][j] = fmaT(cv[8 + 2 * j + 1], w[r].im, fmaT(cv[8 + 2 * j], w[r].re, mi[v][j]));
                }
            }
        }
        T A, D, Br, Bi;
        auto stage_h = [&](int j) {
            A = hs[0]; D = hs[1]; Br = hs[2]; Bi = hs[3];
#pragma unroll
            for (int x = 0; x < NX; ++x) {
                const T f = (T)row[3 * j + x];
                A = fmaT(f, xs[x][0], A); D = fmaT(f, xs[x][1], D);
                Br = fmaT(f, xs[x][2], Br); Bi = fmaT(f, xs[x][3], Bi);
            }
        };
        T kp, kr, ki;
        auto deriv = [&](T sp, T sq, T scr, T sci) {
            kp = (T)2 * fmaT(Bi, scr, -(Br * sci));
            const T ad = A - D, qp = sq - sp;
            kr = fmaT(ad, sci, qp * Bi);
            ki = -fmaT(ad, scr, qp * Br);
        };
        const T P = (T)p, Q = (T)q, Cr = (T)c.re, Ci = (T)c.im;
        stage_h(0);
        deriv(P, Q, Cr, Ci);
        T ap = kp, ar = kr, ai = ki;
        stage_h(1);
        deriv(fmaT(half, kp, P), fmaT(-half, kp, Q), fmaT(half, kr, Cr), fmaT(half, ki, Ci));
        ap = fmaT((T)2, kp, ap); ar = fmaT((T)2, kr, ar); ai = fmaT((T)2, ki, ai);
        deriv(fmaT(half, kp, P), fmaT(-half, kp, Q), fmaT(half, kr, Cr), fmaT(half, ki, Ci));
        ap = fmaT((T)2, kp, ap); ar = fmaT((T)2, kr, ar); ai = fmaT((T)2, ki, ai);
        stage_h(2);
        deriv(fmaT(full, kp, P), fmaT(-full, kp, Q), fmaT(full, kr, Cr), fmaT(full, ki, Ci));
        ap += kp; ar += kr; ai += ki;
        p = fmaT(pref, (double)ap, p);
        q = fmaT(-pref, (double)ap, q);
        c.re = fmaT(pref, (double)ar, c.re);
        c.im = fmaT(pref, (double)ai, c.im);
        double jc[C];
#pragma unroll
        for (int v = 0; v < NV; ++v) {  // inactive lanes hold rho = 0 and contribute 0
            jc[2 * v] = fmaT(mr[v][3], c.im, fmaT(mr[v][2], c.re, fmaT(mr[v][1], q, mr[v][0] * p)));
            jc[2 * v + 1] = fmaT(mi[v][3], c.im, fmaT(mi[v][2], c.re, fmaT(mi[v][1], q, mi[v][0] * p)));
        }
        JS::put(jl, lane, ti % TCH, jc);
        if (chk && (ti % TCH == TCH - 1 || ti == d.nt - 1))
            JS::flush(jl, lane, ti - ti % TCH, ti % TCH + 1, jpart, nwaves, wave);
    };
    // whole chunks of TCH steps with no branch inside (unrolled by two up to
    // 12 pairs; 32 pairs would spill), then the rest step by step: the current
    // of one step and the next step's sums and chain are scheduled together
    constexpr int UNR = NREP <= 12 ? 2 : 1;
    int ti = 0;
    for (; ti + TCH <= d.nt; ti += TCH) {
#pragma unroll UNR
        for (int u = 0; u < TCH; ++u) step(ti + u, false);
        JS::flush(jl, lane, ti, TCH, jpart, nwaves, wave);
    }
    for (; ti < d.nt; ++ti) step(ti, true);
    if (active) {
        double* o = rho_out + 2 * (size_t)k;
        const size_t st = 2 * (size_t)d.N;
        o[0] = p; o[1] = 0.0;
        o[st] = c.re; o[st + 1] = c.im;
        o[2 * st] = c.re; o[2 * st + 1] = -c.im;
        o[3 * st] = q; o[3 * st + 1] = 0.0;
    }
}

// nb = 2 in Bloch-vector form with the pair coefficients streamed (complex
// window current): the streamed counterpart of bloch2_body for models whose
// every k-point has the same occupation. Per pair and step: the twiddled phase
// w, 3 Fourier-sum components for g and for each rbc slot, 6 for each current
// axis, instead of 4 and 8 in pair2s_body; a stage is one cross product. The
// r0 (m_0 + m_1) part of the current sums, over the grid, to the constant of
// the origin, which the reduction adds.
template <int NREP, int NX, int NV, typename T>
__device__ __forceinline__ void bloch2s_body(const SolveDesc& d, const PairDesc& pd, const InstPtrs& ip, int bid) {
    const cdptr steps = as_const(ip.psteps);
    const double* __restrict__ rho_in = ip.rho_in;
    double* __restrict__ rho_out = ip.rho_out;
    double* __restrict__ jpart = ip.jpart;
    const int nwaves = ip.nwaves;
    const double dt = ip.dt;
    const int k = bid * 64 + threadIdx.x;
    const bool active = k < d.N;
    const int kk = active ? k : 0;
    const int m1 = kk / d.N2, m2 = kk - (kk / d.N2) * d.N2;
    const int wave = bid;
    const int lane = threadIdx.x & 63;
    constexpr int C = 2 * NV;
    constexpr int TCH = 64 / C;
    using JS = JStage<C, TCH>;
    __shared__ double jl[64 * JS::ROW];
    constexpr int NXA = NX > 0 ? NX : 1;

    cx<double> w0[NREP];
#pragma unroll
    for (int r = 0; r < NREP; ++r) {
        w0[r] = {0.0, 0.0};
        if (r < pd.nrep) {
            const int j1 = posmod((long)m1 * pd.n1[r], d.N1), j2 = posmod((long)m2 * pd.n2[r], d.N2);
            w0[r] = cmul(ld<double>(ip.tw1 + 2 * j1), ld<double>(ip.tw2 + 2 * j2));
        }
    }
    double r0 = 0, sx = 0, sy = 0, sz = 0;
    if (active) {
        const double p = rho_in[2 * (size_t)k];
        const cx<double> c = ld<double>(rho_in + 1 * 2 * (size_t)d.N + 2 * k);
        const double q = rho_in[3 * 2 * (size_t)d.N + 2 * k];
        r0 = 0.5 * (p + q);
        sz = 0.5 * (p - q);
        sx = c.re;
        sy = -c.im;
    }
    const T half = (T)(0.5 * dt);
    const T full = (T)dt;
    const double pref = 0.1666666666666 * dt;

    auto step = [&](int ti, bool chk) {
        const cdptr row = steps + (size_t)ti * pd.stride;
        const cdptr cf = as_const(ip.scoef) + BS_OFF + opaque_zero();
        const cfptr cf32 = (cfptr)(as_const(ip.scoef) + BS_OFF + BS_N) + opaque_zero();
        auto cT = [&](int i) -> T {  // g and rbc coefficient i in the stage type
            if constexpr (std::is_same<T, float>::value) return cf32[i];
            else return (T)cf[i];
        };
        cx<double> w[NREP];
#pragma unroll
        for (int r = 0; r < NREP; ++r) w[r] = cmul(w0[r], ld<double>(row + PAIR_STEP_HEAD + 2 * r));
        T gs[3], xs[NXA][3];
        double ur[NV][3], ui[NV][3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            gs[j] = cT(j);
#pragma unroll
            for (int x = 0; x < NX; ++x) xs[x][j] = cT(3 + 3 * x + j);
#pragma unroll
            for (int v = 0; v < NV; ++v) { ur[v][j] = cf[12 + 3 * v + j]; ui[v][j] = cf[21 + 3 * v + j]; }
        }
#pragma unroll
        for (int r = 0; r < NREP; ++r) {
            const int o = BS_HEAD + BS_REP * r;
            const T wr = (T)w[r].re, wi = (T)w[r].im;
#pragma unroll
            for (int j = 0; j < 3; ++j) gs[j] = fmaT(cT(o + 2 * j + 1), wi, fmaT(cT(o + 2 * j), wr, gs[j]));
#pragma unroll
            for (int x = 0; x < NX; ++x)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    xs[x][j] = fmaT(cT(o + 6 + 6 * x + 2 * j + 1), wi, fmaT(cT(o + 6 + 6 * x + 2 * j), wr, xs[x][j]));
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const cdptr cv = cf + o + 24 + 12 * v;
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    ur[v][j] = fmaT(cv[2 * j + 1], w[r].im, fmaT(cv[2 * j], w[r].re, ur[v][j]));
                    ui[v][j] = fmaT(cv[6 + 2 * j + 1], w[r].im, fmaT(cv[6 + 2 * j], w[r].re, ui[v][j]));
                }
            }
        }
        T gx, gy, gz;
        auto stage_g = [&](int j) {
            gx = gs[0]; gy = gs[1]; gz = gs[2];
#pragma unroll
            for (int x = 0; x < NX; ++x) {
                const T f = (T)row[3 * j + x];
                gx = fmaT(f, xs[x][0], gx); gy = fmaT(f, xs[x][1], gy); gz = fmaT(f, xs[x][2], gz);
            }
        };
        T kx, ky, kz;
        auto deriv = [&](T ax, T ay, T az) {  // k = g x s
            kx = fmaT(gy, az, -(gz * ay));
            ky = fmaT(gz, ax, -(gx * az));
            kz = fmaT(gx, ay, -(gy * ax));
        };
        const T Sx = (T)sx, Sy = (T)sy, Sz = (T)sz;
        stage_g(0);
        deriv(Sx, Sy, Sz);
        T ax = kx, ay = ky, az = kz;
        stage_g(1);
        deriv(fmaT(half, kx, Sx), fmaT(half, ky, Sy), fmaT(half, kz, Sz));
        ax = fmaT((T)2, kx, ax); ay = fmaT((T)2, ky, ay); az = fmaT((T)2, kz, az);
        deriv(fmaT(half, kx, Sx), fmaT(half, ky, Sy), fmaT(half, kz, Sz));
        ax = fmaT((T)2, kx, ax); ay = fmaT((T)2, ky, ay); az = fmaT((T)2, kz, az);
        stage_g(2);
        deriv(fmaT(full, kx, Sx), fmaT(full, ky, Sy), fmaT(full, kz, Sz));
        ax += kx; ay += ky; az += kz;
        sx = fmaT(pref, (double)ax, sx);
        sy = fmaT(pref, (double)ay, sy);
        sz = fmaT(pref, (double)az, sz);
        // current of the new state (FP64), without the r0 term; inactive lanes
        // hold s = 0 and contribute exactly 0
        double jc[C];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            jc[2 * v] = fmaT(ur[v][2], sz, fmaT(ur[v][1], sy, ur[v][0] * sx));
            jc[2 * v + 1] = fmaT(ui[v][2], sz, fmaT(ui[v][1], sy, ui[v][0] * sx));
        }
        JS::put(jl, lane, ti % TCH, jc);
        if (chk && (ti % TCH == TCH - 1 || ti == d.nt - 1))
            JS::flush(jl, lane, ti - ti % TCH, ti % TCH + 1, jpart, nwaves, wave);
    };
    int ti = 0;
    for (; ti + TCH <= d.nt; ti += TCH) {
#pragma unroll 2
        for (int u = 0; u < TCH; ++u) step(ti + u, false);
        JS::flush(jl, lane, ti, TCH, jpart, nwaves, wave);
    }
    for (; ti < d.nt; ++ti) step(ti, true);
    if (active) {
        double* o = rho_out + 2 * (size_t)k;
        const size_t st = 2 * (size_t)d.N;
        o[0] = r0 + sz; o[1] = 0.0;
        o[st] = sx; o[st + 1] = -sy;
        o[2 * st] = sx; o[2 * st + 1] = sy;
        o[3 * st] = r0 - sz; o[3 * st + 1] = 0.0;
    }
}

// --------------------------------------------------------------------------
// Element-per-lane kernel for any nb <= 64. A group of G lanes owns one k
// (G = nb*nb rounded up to a power of two, at most 64, for nb <= 16; the whole
// 256-lane workgroup for 17 <= nb <= 32, a 1024-lane one for 33-64); lane l
// owns the matrix elements e = l + G*q, q < EPL (1 for nb <= 8, 4 for 9-16,
// 2 while nb*nb <= 512, else 4). Per stage the stage Hamiltonian and stage rho
// are staged through LDS and each lane forms its commutator elements from
// row i / column j there. LDS per k: the support phases (ns complex, FP64) and
// two nb*nb matrices of type T (128 KiB at 64 bands in FP64). BIG (33-64
// bands, 128 VGPRs at 16 waves per workgroup): the stage Hamiltonians are
// summed afresh for each stage field instead of keeping the field-free sums of
// H and rbc in registers.
__host__ __device__ inline int lanes_per_k(int nb) {
    if (nb > 32) return 1024;
    if (nb > 16) return 256;
    int g = 1;
    while (g < nb * nb && g < 64) g <<= 1;
    return g;
}

template <int EPL, typename T, bool BIG = false>
__device__ __forceinline__ void elem_body(const SolveDesc& d, const InstPtrs& ip, int bid) {
    const double* __restrict__ blk = ip.blk;
    const double* __restrict__ steptab = ip.steptab;
    const double* __restrict__ tw1 = ip.tw1;
    const double* __restrict__ tw2 = ip.tw2;
    const double* __restrict__ rho_in = ip.rho_in;
    double* __restrict__ rho_out = ip.rho_out;
    double* __restrict__ jpart = ip.jpart;
    const int nwaves = ip.nwaves;
    const double dt = ip.dt;
    const int G = lanes_per_k(d.nb);
    extern __shared__ __align__(16) unsigned char smem[];
    const int NB = d.nb, E = NB * NB;
    const int kpb = blockDim.x / G;
    const int grp = threadIdx.x / G;
    const int l = threadIdx.x - grp * G;
    const int k = bid * kpb + grp;
    const bool kvalid = k < d.N;
    const int kk = kvalid ? k : 0;
    const int m1 = kk / d.N2, m2 = kk - (kk / d.N2) * d.N2;
    const int wave = (bid * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;

    cx<double>* wS = reinterpret_cast<cx<double>*>(smem) + (size_t)grp * d.ns;
    cx<T>* lH = reinterpret_cast<cx<T>*>(reinterpret_cast<cx<double>*>(smem) + (size_t)kpb * d.ns) + (size_t)grp * 2 * E;
    cx<T>* lR = lH + E;

    int ei[EPL], eI[EPL], eJ[EPL];
    bool ev[EPL];
#pragma unroll
    for (int q = 0; q < EPL; ++q) {
        const int e = l + G * q;
        ev[q] = kvalid && e < E;
        ei[q] = e < E ? e : 0;
        eI[q] = ei[q] / NB;
        eJ[q] = ei[q] - eI[q] * NB;
    }
    cx<double> rho[EPL];
#pragma unroll
    for (int q = 0; q < EPL; ++q)
        rho[q] = ev[q] ? ld<double>(rho_in + 2 * ((size_t)ei[q] * d.N + k)) : cx<double>{0.0, 0.0};

    const T half = (T)(0.5 * dt);
    const T full = (T)dt;
    const double pref = 0.1666666666666 * dt;

    for (int ti = 0; ti < d.nt; ++ti) {
        const double* st = steptab + (size_t)ti * d.stride;
        for (int s = l; s < d.ns; s += G) {
            const int j1 = posmod((long)m1 * d.n1[s], d.N1), j2 = posmod((long)m2 * d.n2[s], d.N2);
            wS[s] = cmul(cmul(ld<double>(tw1 + 2 * j1), ld<double>(tw2 + 2 * j2)), ld<double>(st + STEP_HEAD + 2 * s));
        }
        __syncthreads();
        constexpr int KEEP = BIG ? 1 : EPL;  // field-free sums kept in registers (not for BIG)
        cx<T> Ht[KEEP], Xt[3][KEEP];
        if constexpr (!BIG) {
#pragma unroll
            for (int q = 0; q < EPL; ++q) {
                Ht[q] = {0, 0};
                Xt[0][q] = Xt[1][q] = Xt[2][q] = {0, 0};
            }
            for (int s = 0; s < d.ns; ++s) {
                const uint32_t om = d.opmask[s];
                const cx<T> w = cast<T>(wS[s]);
                const double* b = blk + (size_t)s * NOPS * E * 2;
                if (om & (1u << OP_H)) {
#pragma unroll
                    for (int q = 0; q < EPL; ++q) cmac(Ht[q], w, ld<T>(b + 2 * (OP_H * E + ei[q])));
                }
                for (int a = 0; a < 3; ++a)
                    if (om & (1u << (OP_X + a))) {
#pragma unroll
                        for (int q = 0; q < EPL; ++q) cmac(Xt[a][q], w, ld<T>(b + 2 * ((OP_X + a) * E + ei[q])));
                    }
            }
        }
        cx<T> Hs[EPL], rs[EPL], K[EPL], acc[EPL];
        auto stage_h = [&](int js) {
            if constexpr (BIG) {
                // sum_s w_s (H_s + E_js . rbc_s), in the same operation order as
                // the kept sums below: H and each rbc axis summed over s first;
                // one element at a time (register pressure)
#pragma unroll 1
                for (int q = 0; q < EPL; ++q) {
                    cx<T> h = {0, 0}, x[3] = {{0, 0}, {0, 0}, {0, 0}};
                    for (int s = 0; s < d.ns; ++s) {
                        const uint32_t om = d.opmask[s];
                        const cx<T> w = cast<T>(wS[s]);
                        const double* b = blk + (size_t)s * NOPS * E * 2;
                        if (om & (1u << OP_H)) cmac(h, w, ld<T>(b + 2 * (OP_H * E + ei[q])));
                        for (int a = 0; a < 3; ++a)
                            if (om & (1u << (OP_X + a))) cmac(x[a], w, ld<T>(b + 2 * ((OP_X + a) * E + ei[q])));
                    }
                    for (int a = 0; a < 3; ++a)
                        if (d.xmask & (1 << a)) {
                            const T f = (T)st[3 * js + a];
                            h.re += f * x[a].re;
                            h.im += f * x[a].im;
                        }
                    Hs[q] = h;
                }
            } else {
#pragma unroll
                for (int q = 0; q < EPL; ++q) {
                    cx<T> h = Ht[q];
                    for (int a = 0; a < 3; ++a)
                        if (d.xmask & (1 << a)) {
                            const T f = (T)st[3 * js + a];
                            h.re += f * Xt[a][q].re;
                            h.im += f * Xt[a][q].im;
                        }
                    Hs[q] = h;
                }
            }
        };
        auto deriv = [&](bool new_h) {  // K = -i (Hs rs - rs Hs), element-wise
#pragma unroll
            for (int q = 0; q < EPL; ++q)
                if (l + G * q < E) {
                    if (new_h) lH[ei[q]] = Hs[q];
                    lR[ei[q]] = rs[q];
                }
            __syncthreads();
            if (d.hermhx) {
                // Hermitian H and rho: K = -i (M - M^dagger) with M = H rho, one
                // product; M replaces rho in LDS for the transposed reads
                cx<T> M[EPL];
#pragma unroll
                for (int q = 0; q < EPL; ++q) {
                    cx<T> c = {0, 0};
                    const cx<T>* hrow = lH + eI[q] * NB;
#pragma unroll 4
                    for (int m = 0; m < NB; ++m) cmac(c, hrow[m], lR[m * NB + eJ[q]]);
                    M[q] = c;
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < EPL; ++q)
                    if (l + G * q < E) lR[ei[q]] = M[q];
                __syncthreads();
#pragma unroll
                for (int q = 0; q < EPL; ++q) {
                    const cx<T> t = lR[eJ[q] * NB + eI[q]];  // M_ji
                    K[q] = {M[q].im + t.im, t.re - M[q].re};
                }
            } else {
#pragma unroll
                for (int q = 0; q < EPL; ++q) {
                    cx<T> c = {0, 0};
                    const cx<T>* hrow = lH + eI[q] * NB;
                    const cx<T>* rrow = lR + eI[q] * NB;
#pragma unroll 4
                    for (int m = 0; m < NB; ++m) {
                        cmac(c, hrow[m], lR[m * NB + eJ[q]]);
                        cmsub(c, rrow[m], lH[m * NB + eJ[q]]);
                    }
                    K[q] = {c.im, -c.re};
                }
            }
            __syncthreads();
        };
        stage_h(0);
#pragma unroll
        for (int q = 0; q < EPL; ++q) rs[q] = cast<T>(rho[q]);
        deriv(true);
#pragma unroll
        for (int q = 0; q < EPL; ++q) {
            acc[q] = K[q];
            rs[q] = {(T)rho[q].re + half * K[q].re, (T)rho[q].im + half * K[q].im};
        }
        stage_h(1);
        deriv(true);
#pragma unroll
        for (int q = 0; q < EPL; ++q) {
            acc[q].re += (T)2 * K[q].re; acc[q].im += (T)2 * K[q].im;
            rs[q] = {(T)rho[q].re + half * K[q].re, (T)rho[q].im + half * K[q].im};
        }
        deriv(false);
#pragma unroll
        for (int q = 0; q < EPL; ++q) {
            acc[q].re += (T)2 * K[q].re; acc[q].im += (T)2 * K[q].im;
            rs[q] = {(T)rho[q].re + full * K[q].re, (T)rho[q].im + full * K[q].im};
        }
        stage_h(2);
        deriv(true);
#pragma unroll
        for (int q = 0; q < EPL; ++q) {
            acc[q].re += K[q].re; acc[q].im += K[q].im;
            if (ev[q]) {
                rho[q].re += pref * (double)acc[q].re;
                rho[q].im += pref * (double)acc[q].im;
            }
        }
        double jc[6] = {0, 0, 0, 0, 0, 0};
        for (int s = 0; s < d.ns; ++s) {
            const uint32_t om = d.opmask[s];
            if (!(om & (7u << OP_V))) continue;
            const cx<double> w = wS[s];
            const double* b = blk + (size_t)s * NOPS * E * 2;
            for (int a = 0; a < 3; ++a)
                if (om & (1u << (OP_V + a))) {
                    cx<double> part = {0.0, 0.0};
#pragma unroll
                    for (int q = 0; q < EPL; ++q)
                        if (ev[q]) cmac(part, ld<double>(b + 2 * ((OP_V + a) * E + ei[q])), cx<double>{rho[q].re, -rho[q].im});
                    const cx<double> c = cmul(part, w);
                    jc[2 * a] += c.re;
                    jc[2 * a + 1] += c.im;
                }
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) jc[c] = wave_sum(jc[c]);
        if (lane == 0) {
            double* out = jpart + ((size_t)ti * nwaves + wave) * 6;
#pragma unroll
            for (int c = 0; c < 6; ++c) out[c] = jc[c];
        }
        __syncthreads();  // wS is rewritten by the next step
    }
#pragma unroll
    for (int q = 0; q < EPL; ++q)
        if (ev[q]) {
            rho_out[2 * ((size_t)ei[q] * d.N + k)] = rho[q].re;
            rho_out[2 * ((size_t)ei[q] * d.N + k) + 1] = rho[q].im;
        }
}

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// --------------------------------------------------------------------------
// Row-separable element kernel: Hermitian H and rbc, 3 <= nb <= NBP (4, 8 or
// 16, smaller nb zero-padded). A workgroup of W waves owns a chunk of one k-row
// m1. Per step it builds, cooperatively and double-buffered in LDS,
//   G[j][d]  = sum_{s: n2_s = n2_d} tw1[m1 n1_s] e^{iA.R_s} (H_s + E_j . rbc_s)   (stage fields E(t_i), E(t_i+1))
//   GV[d][a] = sum_{s: n2_s = n2_d} tw1[m1 n1_s] e^{iA.R_s} v_a,s                 (axes a = x, y, z)
// (for NBP = 16 GV holds only the values d with a velocity block, by rank)
// and every k-point forms its stage Hamiltonian sum_d tw2[m2 n2_d] G[j][d]
// with MAXD complex MACs per element instead of ns (the midpoint stages'
// field is the mean of the two, and so is their Hamiltonian). For Hermitian H
// and X the commutator is -i(M - M^dagger) with M = H X: one complex product
// per stage. Each lane owns a 2x2 element block, so every LDS read of the
// product feeds two MACs; the (NBP/2)^2 lanes of a k-point sit in one wave and
// exchange H, X and M through LDS with wave-level ordering only (4 bands in
// FP32: a column per lane and the commutator from registers on the matrix
// cores, below). Each lane group carries
// KPW = 1, 2 or 4 k-points of the row (W groups apart) through every step, so
// one build serves W * GPW * KPW k-points; their states rotate through one
// register slot and the current reads GV once for all of them. One workgroup
// barrier per step: the tables of step t+1 are built while step t runs (NBP <= 8). The
// build reads the per-support weights and field row from LDS (written two
// steps ahead) and, on the FP32 path, the H and rbc blocks staged in LDS at
// launch; only the velocity blocks of the window supports come from HBM.
struct SepLayout {
    size_t off_g, gbytes, off_gv, gvbytes, off_sup, off_w, off_b, off_k, kbytes, off_z, off_jw, bytes;
    int nsw;  // support slots of the per-support tables (ns rounded up to 4)
};
// Tables of NBP <= 8 are double-buffered (built one step ahead); NBP = 16
// keeps one buffer and builds between two barriers.
__host__ __device__ constexpr int sep_nbuf(int NBP) { return NBP <= 8 ? 2 : 1; }
// Most k-points a wave carries through a step under one table build (the
// launch picks 1 .. this, a power of two)
__host__ __device__ constexpr int sep_kpw_max(int NBP) { return NBP == 16 ? 4 : NBP == 8 ? 2 : 1; }
// 3-4 bands in FP32: the commutator on the matrix cores (v_mfma_f32_4x4x1f32,
// 16 blocks of 4 lanes). Each lane owns one column of its k-point (4 lanes =
// one block) and HX - XH is formed from registers, with no LDS exchange.
__host__ __device__ constexpr bool sep_mf(int NBP, int tsize) { return NBP == 4 && tsize == 4; }
// nstage: support blocks x elements whose H and rbc blocks are staged in LDS
// (FP32 path, NBP <= 8: ns nb^2), 0 when the table build reads them from HBM.
// nwd: values d with a velocity block; for NBP = 16, where the LDS bounds the
// waves per workgroup, the velocity tables GV hold only those (by rank)
__host__ __device__ inline SepLayout sep_layout(int NBP, int MAXD, int tsize, int W, size_t nstage, int nwd, int ns) {
    SepLayout L;
    L.nsw = (ns + 3) & ~3;
    const size_t E = (size_t)NBP * NBP;
    const int GPW = 64 / ((NBP / 2) * (NBP / 2));
    const int nbuf = sep_nbuf(NBP);
    const bool mf = sep_mf(NBP, tsize);
    size_t o = 0;
    L.gbytes = 2 * (size_t)MAXD * E * 2 * tsize;  // fields E(t_i) and E(t_i+1)
    L.off_g = o; o = align16(o + nbuf * L.gbytes);
    L.gvbytes = (size_t)(NBP == 16 ? nwd : MAXD) * 3 * E * 16;
    L.off_gv = o; o = align16(o + nbuf * L.gvbytes);
    L.off_sup = o; o = align16(o + (size_t)L.nsw * 16 + (size_t)L.nsw * 4 + 16 * 4 + (size_t)MAXD * E * 2);
    if (nbuf == 2) { L.off_w = o; o = align16(o + 2 * ((size_t)L.nsw * 16 + 16 * 8)); }  // weights + field rows
    else L.off_w = 0;
    L.off_b = o; o = align16(o + nstage * 4 * 8);
    // H/M and X of one k-point; the pad staggers banks (none in the matrix-core
    // form, whose commutator stays in registers)
    L.kbytes = mf ? 0 : 2 * E * 2 * tsize + 64;
    L.off_k = o; o = align16(o + (size_t)W * GPW * L.kbytes);
    // row twiddles of the k-points of each wave when a wave may carry several
    const int kmax = sep_kpw_max(NBP);
    L.off_z = o; o = align16(o + (kmax > 1 ? (size_t)W * kmax * GPW * MAXD * 16 : 0));
    L.off_jw = o; o = align16(o + sizeof(double) * 2 * W * 6);
    L.bytes = o;
    return L;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// Orders the LDS accesses of one wave: writes before the barrier are seen by
// every lane's reads after it (LDS executes a wave's instructions in order).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Element placement of an NBP x NBP tile in LDS, chosen so that the 16 lanes
// of a k-point (2x2 blocks (bi, bj)) touch one contiguous 256-byte span in
// every access: block-row writes, row pairs (m, i0..i0+1) of the product and
// the transposed pairs. FP32: planes by row parity, then block, then the
// column pair (16 bytes); FP64: planes by (row, column) parity, then block.
template <typename T, int NBP>
struct Tile;
template <int NBP>
struct Tile<float, NBP> {
    static constexpr int HB = NBP / 2, B2 = HB * HB;
    __host__ __device__ static int off(int i, int j) { return (((i & 1) * B2 + (i >> 1) * HB + (j >> 1)) << 1) + (j & 1); }
    __device__ static void ld2(const cx<float>* t, int i, int j0, cx<float>& a, cx<float>& b) {
        const float4 v = *reinterpret_cast<const float4*>(t + off(i, j0));
        a = {v.x, v.y};
        b = {v.z, v.w};
    }
    __device__ static void st2(cx<float>* t, int i, int j0, cx<float> a, cx<float> b) {
        *reinterpret_cast<float4*>(t + off(i, j0)) = make_float4(a.re, a.im, b.re, b.im);
    }
};
template <int NBP>
struct Tile<double, NBP> {
    static constexpr int HB = NBP / 2, B2 = HB * HB;
    __host__ __device__ static int off(int i, int j) { return (((i & 1) << 1) + (j & 1)) * B2 + (i >> 1) * HB + (j >> 1); }
    __device__ static void ld2(const cx<double>* t, int i, int j0, cx<double>& a, cx<double>& b) {
        const double2 v0 = *reinterpret_cast<const double2*>(t + off(i, j0));
        const double2 v1 = *reinterpret_cast<const double2*>(t + off(i, j0) + B2);
        a = {v0.x, v0.y};
        b = {v1.x, v1.y};
    }
    __device__ static void st2(cx<double>* t, int i, int j0, cx<double> a, cx<double> b) {
        *reinterpret_cast<double2*>(t + off(i, j0)) = make_double2(a.re, a.im);
        *reinterpret_cast<double2*>(t + off(i, j0) + B2) = make_double2(b.re, b.im);
    }
};

// Tiles of the 4-band matrix-core form, where each lane owns one column of
// its k-point. FP32 (stage tables): 16-byte chunks of two rows (i, i + 1) of a
// column, laid out [row pair][column]; FP64 (velocity tables): row-major.
template <typename T, int NBP>
struct SegTile {
    __host__ __device__ static int off(int i, int j) {
        if constexpr (std::is_same<T, float>::value) return 2 * (NBP * (i >> 1) + j) + (i & 1);
        else return i * NBP + j;
    }
    // rows i0 .. i0 + 3 of column j (i0 even)
    __device__ static void ld4(const cx<T>* t, int i0, int j, cx<T> (&v)[4]) {
        if constexpr (std::is_same<T, float>::value) {
#pragma unroll
            for (int q = 0; q < 4; q += 2) {
                const float4 a = *reinterpret_cast<const float4*>(t + off(i0 + q, j));
                v[q] = {a.x, a.y};
                v[q + 1] = {a.z, a.w};
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double2 a = *reinterpret_cast<const double2*>(t + off(i0 + q, j));
                v[q] = {a.x, a.y};
            }
        }
    }
};

template <typename T>
__device__ __forceinline__ void cfma(cx<T>& acc, cx<T> a, cx<T> b) {  // acc += a*b
    acc.re = fmaT(a.re, b.re, fmaT(-a.im, b.im, acc.re));
    acc.im = fmaT(a.re, b.im, fmaT(a.im, b.re, acc.im));
}
template <typename T>
__device__ __forceinline__ void cfmac(cx<T>& acc, cx<T> a, cx<T> b) {  // acc += conj(a)*b
    acc.re = fmaT(a.re, b.re, fmaT(a.im, b.im, acc.re));
    acc.im = fmaT(a.re, b.im, fmaT(-a.im, b.re, acc.im));
}

// Sum over the 16 lanes of each DPP row, returned in every lane of the row
// (quad swaps, half-row mirror, row mirror); no LDS traffic.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double row_sum16(double v) {
    v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_mov<0x141>(v);  // row_half_mirror
    v += dpp_mov<0x140>(v);  // row_mirror
    return v;
}
// Wave total, valid in every lane: row sums combined through readlane.
__device__ __forceinline__ double wave_total(double v) {
    v = row_sum16(v);
    auto rl = [&](int l) {
        return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                                __builtin_amdgcn_readlane(__double2loint(v), l));
    };
    return (rl(0) + rl(16)) + (rl(32) + rl(48));
}

// LEAN: the FP32 form of 5-8 bands within 168 VGPRs (3 waves per SIMD):
// no LDS staging of the blocks (the H and rbc blocks are rounded to FP32 as
// they are read, so the tables equal the staged form's) and the field row
// through scalar loads.
template <int NBP, int MAXD, typename T, bool LEAN = false>
__device__ __forceinline__ void sep_body(const SolveDesc& d, const InstPtrs& ip, int tile) {
    constexpr int HB = NBP / 2, LPK = HB * HB, GPW = 64 / LPK, E = NBP * NBP;
    // MF: the 4-band FP32 matrix-core form, a column per lane
    constexpr bool MF = sep_mf(NBP, (int)sizeof(T));
    using TT = typename std::conditional<MF, SegTile<T, NBP>, Tile<T, NBP>>::type;
    using TD = typename std::conditional<MF, SegTile<double, NBP>, Tile<double, NBP>>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    const int W = blockDim.x >> 6;
    // FP32 path, NBP <= 8: the H and rbc blocks are staged once in LDS as FP32
    constexpr bool STAGE = std::is_same<T, float>::value && NBP <= 8 && !LEAN;
    const SepLayout L = sep_layout(NBP, MAXD, (int)sizeof(T), W, STAGE ? (size_t)d.ns * d.nb * d.nb : 0, d.nwd, d.ns);
    const int nb = d.nb, nbe = nb * nb, nd = d.nd, N = d.N, N1 = d.N1, N2 = d.N2;
    const int tid = threadIdx.x, bd = blockDim.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int g = lane / LPK, bl = lane - g * LPK;
    const int i0 = 2 * (bl / HB), j0 = 2 * (bl - (bl / HB) * HB);
    // element q = 0..3 of the lane: 2x2 block (i0, j0), or (MF) row q of column bl
    auto eli = [&](int q) { return MF ? q : i0 + (q >> 1); };
    auto elj = [&](int q) { return MF ? bl : j0 + (q & 1); };
    // KPW k-points per lane group (ip.kpw, 1 .. sep_kpw_max(NBP)) share each
    // table build; k-point p of group g of wave w sits at m2 = base + (p W + w) GPW + g
    constexpr int KPW_MAX = sep_kpw_max(NBP);
    const int KPW = KPW_MAX > 1 ? ip.kpw : 1;
    const int kpt = W * GPW * KPW;
    const int tpr = (N2 + kpt - 1) / kpt;
    const int m1 = tile / tpr;
    auto m2_at = [&](int p, int gg) { return (tile - m1 * tpr) * kpt + (p * W + wave) * GPW + gg; };
    auto m2_of = [&](int p) { return m2_at(p, g); };
    cx<T>* bufH = reinterpret_cast<cx<T>*>(smem + L.off_k + (size_t)(wave * GPW + g) * L.kbytes);  // H, then M
    cx<T>* bufX = bufH + E;
    double* jw = reinterpret_cast<double*>(smem + L.off_jw);
    // per-support table in n2-sorted order: row twiddle tw1[m1 n1_s], index s;
    // segment bounds of each value d; entry codes (d << 16 | el << 8 | i << 4 | j)
    cx<double>* sw1 = reinterpret_cast<cx<double>*>(smem + L.off_sup);
    int* ss = reinterpret_cast<int*>(smem + L.off_sup + (size_t)L.nsw * 16);
    int* seg = ss + L.nsw;
    uint16_t* ent = reinterpret_cast<uint16_t*>(seg + 16);
    const int nent = nd * nbe;

    for (size_t i = tid; i < (L.off_sup - L.off_g) / 16; i += bd)
        reinterpret_cast<double2*>(smem + L.off_g)[i] = make_double2(0.0, 0.0);
    // ss[q]: support index | 1 << 16 when it has a velocity block
    for (int q = tid; q < d.ns; q += bd) {
        const int s = d.sorder[q];
        ss[q] = s | ((d.opmask[s] & (7u << OP_V)) ? 1 << 16 : 0);
        sw1[q] = ld<double>(ip.tw1 + 2 * posmod((long)m1 * d.n1[s], N1));
    }
    for (int q = tid; q <= nd; q += bd) seg[q] = d.dseg[q];
    float2* bst = reinterpret_cast<float2*>(smem + L.off_b);  // [n2-sorted q][H, rbc x, y, z][nb^2]
    if constexpr (STAGE) {
        for (int i = tid; i < d.ns * 4 * nbe; i += bd) {
            const int q = i / (4 * nbe), r = i - q * 4 * nbe;  // r = op * nbe + el, op 0..3 = H, rbc
            const double* b = ip.blk + 2 * ((size_t)d.sorder[q] * NOPS * nbe + r);
            bst[i] = make_float2((float)b[0], (float)b[1]);
        }
    }
    // Double-buffered tables (NBP <= 8): the build of step ti + 1 runs during
    // step ti, and its per-support weights tw1[m1 n1_s] e^{i A.R_s} (n2-sorted)
    // and field row are written two steps ahead into slot ti & 1, so the build
    // reads LDS only (and the velocity blocks). Single buffer (NBP = 16): the
    // build reads the phases from the step table itself.
    constexpr bool DB = sep_nbuf(NBP) == 2;
    cx<double>* wv = reinterpret_cast<cx<double>*>(smem + L.off_w);
    double* efs = reinterpret_cast<double*>(smem + L.off_w + 2 * (size_t)L.nsw * 16);
    auto weights = [&](int ti) {
        const double* st = ip.steptab + (size_t)ti * d.stride;
        const int sl = ti & 1;
        for (int q = tid; q < d.ns; q += bd)
            wv[sl * L.nsw + q] = cmul(sw1[q], ld<double>(st + STEP_HEAD + 2 * (ss[q] & 0xffff)));
        for (int c = tid; c < 9; c += bd) efs[sl * 16 + c] = st[c];
    };
    for (int q = tid; q < nent; q += bd) {
        const int dd = q / nbe, el = q - dd * nbe;
        ent[q] = (uint16_t)((dd << 8) | ((el / nb) << 4) | (el - (el / nb) * nb));  // dd < 8, i, j < 16
    }
    __syncthreads();

    int wmask = 0;  // values d with a velocity block
#pragma unroll
    for (int dd = 0; dd < MAXD; ++dd)
        if (dd < nd && d.wdi[dd] >= 0) wmask |= 1 << dd;
    // GV slot of value d: its rank among those for NBP = 16, d itself otherwise
    auto wrank = [&](int dd) { return NBP == 16 ? __popc((unsigned)wmask & ((1u << dd) - 1u)) : dd; };
    auto build = [&](int ti, int buf) {
        const double* st = ip.steptab + (size_t)ti * d.stride;
        double ef[9];
        if constexpr (LEAN) {  // wave-uniform values in SGPRs
            const cdptr sts = as_const(st);
#pragma unroll
            for (int c = 0; c < 9; ++c) ef[c] = sts[c];
        } else {
#pragma unroll
            for (int c = 0; c < 9; ++c) ef[c] = DB ? efs[(ti & 1) * 16 + c] : st[c];
        }
        const cx<double>* wq = wv + (ti & 1) * L.nsw;
        cx<T>* G = reinterpret_cast<cx<T>*>(smem + L.off_g + buf * L.gbytes);
        cx<double>* GV = reinterpret_cast<cx<double>*>(smem + L.off_gv + buf * L.gvbytes);
        for (int q = tid; q < nent; q += bd) {
            const int code = ent[q];
            const int dd = code >> 8, ei = (code >> 4) & 15, ej = code & 15, el = ei * nb + ej;
            // sum_s w_s B_s for H, the three rbc axes and the three v axes
            cx<double> a[NOPS];
#pragma unroll
            for (int o = 0; o < NOPS; ++o) a[o] = {0, 0};
            const int qe = seg[dd + 1];
#pragma unroll 2  // two supports' block loads in flight (1-4 % at 16 bands, same results)
            for (int qq = seg[dd]; qq < qe; ++qq) {
                const int sv = ss[qq];
                const double* bs = ip.blk + 2 * ((size_t)(sv & 0xffff) * NOPS * nbe + el);
                cx<double> ob[NOPS];
                if constexpr (STAGE) {
                    const float2* b = bst + (size_t)qq * 4 * nbe + el;
#pragma unroll
                    for (int o = 0; o < 4; ++o) ob[o] = {(double)b[o * nbe].x, (double)b[o * nbe].y};
                    if (sv >> 16) {
#pragma unroll
                        for (int o = OP_V; o < NOPS; ++o) ob[o] = ld<double>(bs + 2 * (size_t)o * nbe);
                    } else {
#pragma unroll
                        for (int o = OP_V; o < NOPS; ++o) ob[o] = {0.0, 0.0};
                    }
                } else if constexpr (LEAN) {
                    // the H and rbc blocks rounded to FP32 as the staged form
                    // holds them, so both forms build identical tables
#pragma unroll
                    for (int o = 0; o < OP_V; ++o) {
                        const cx<double> b = ld<double>(bs + 2 * (size_t)o * nbe);
                        ob[o] = {(double)(float)b.re, (double)(float)b.im};
                    }
#pragma unroll
                    for (int o = OP_V; o < NOPS; ++o) ob[o] = ld<double>(bs + 2 * (size_t)o * nbe);
                } else {
#pragma unroll
                    for (int o = 0; o < NOPS; ++o) ob[o] = ld<double>(bs + 2 * (size_t)o * nbe);
                }
                const cx<double> w = DB ? wq[qq] : cmul(sw1[qq], ld<double>(st + STEP_HEAD + 2 * (sv & 0xffff)));
#pragma unroll
                for (int o = 0; o < NOPS; ++o) cfma(a[o], w, ob[o]);
            }
            // stage fields E(t_i) and E(t_i+1) (slots 0, 1); the midpoint stages'
            // H(E2) = (H(E1) + H(E3)) / 2 is formed per k-point, as E2 = (E1 + E3) / 2
#pragma unroll
            for (int jj = 0; jj < 2; ++jj) {
                const int j = 2 * jj;
                cx<double> b = a[OP_H];
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    b.re = fmaT(ef[3 * j + x], a[OP_X + x].re, b.re);
                    b.im = fmaT(ef[3 * j + x], a[OP_X + x].im, b.im);
                }
                G[(size_t)(jj * MAXD + dd) * E + TT::off(ei, ej)] = cast<T>(b);
            }
            if ((wmask >> dd) & 1) {
                const int wr = wrank(dd);
#pragma unroll
                for (int x = 0; x < 3; ++x) GV[(size_t)(wr * 3 + x) * E + TD::off(ei, ej)] = a[OP_V + x];
            }
        }
    };

    cx<double> rhoA[KPW_MAX][4];
#pragma unroll
    for (int p = 0; p < KPW_MAX; ++p) {
        const int m2 = m2_of(p);
        const bool kval = p < KPW && m2 < N2;
        const int k = m1 * N2 + (kval ? m2 : 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int ei = eli(q), ej = elj(q);
            rhoA[p][q] = (kval && ei < nb && ej < nb) ? ld<double>(ip.rho_in + 2 * ((size_t)(ei * nb + ej) * N + k))
                                                      : cx<double>{0.0, 0.0};
        }
    }
    // row twiddles z_d = tw2[m2 n2_d] (0 for k-points past the row end): in
    // registers for one k-point per lane group, else per wave in LDS, [p][g][d]
    // (one address per lane group: broadcasts)
    cx<double> z[KPW_MAX > 1 ? 1 : MAXD];
    cx<double>* zt = reinterpret_cast<cx<double>*>(smem + L.off_z) + (size_t)wave * KPW_MAX * GPW * MAXD;
    if constexpr (KPW_MAX > 1) {
        static_assert(KPW_MAX * GPW * MAXD <= 64, "one lane per twiddle of the wave");
        if (lane < KPW_MAX * GPW * MAXD) {
            const int p = lane / (GPW * MAXD), r = lane - p * GPW * MAXD, gg = r / MAXD, dd = r - gg * MAXD;
            const int m2 = m2_at(p, gg);
            zt[lane] = (p < KPW && dd < nd && m2 < N2) ? ld<double>(ip.tw2 + 2 * posmod((long)m2 * d.dn2[dd], N2))
                                                       : cx<double>{0.0, 0.0};
        }
    } else {
        const int m2 = m2_of(0);
#pragma unroll
        for (int dd = 0; dd < MAXD; ++dd)
            z[dd] = (dd < nd && m2 < N2) ? ld<double>(ip.tw2 + 2 * posmod((long)m2 * d.dn2[dd], N2)) : cx<double>{0.0, 0.0};
    }
    auto zval = [&](int p, int dd) -> cx<double> {
        if constexpr (KPW_MAX > 1) return zt[(p * GPW + g) * MAXD + dd];
        else return z[dd];
    };

    const T half = (T)(0.5 * ip.dt);
    const T full = (T)ip.dt;
    const double pref = 0.1666666666666 * ip.dt;
    if constexpr (DB) {
        weights(0);
        if (d.nt > 1) weights(1);
        __syncthreads();
    }
    build(0, 0);
    __syncthreads();
    for (int ti = 0; ti < d.nt; ++ti) {
        const int buf = DB ? (ti & 1) : 0, jb = ti & 1;
        if (DB && ti + 1 < d.nt) build(ti + 1, buf ^ 1);
        const cx<T>* G = reinterpret_cast<const cx<T>*>(smem + L.off_g + buf * L.gbytes);
        const cx<double>* GV = reinterpret_cast<const cx<double>*>(smem + L.off_gv + buf * L.gvbytes);
        cx<double> jv[3] = {{0, 0}, {0, 0}, {0, 0}};
        // the k-points of the wave in turn; the state sets rotate through
        // rhoA[0] and are back in place after KPW_MAX turns
#pragma unroll 1
        for (int p = 0; p < KPW_MAX; ++p) {
        if (p < KPW) {
        cx<double>* rho = rhoA[0];
        // this k-point's twiddles for the three stage sums: registers for NBP <= 8,
        // LDS broadcasts for NBP = 16 (whose FP64 kernel has no registers left)
        constexpr bool ZREG = KPW_MAX > 1 && NBP != 16;
        cx<T> zr[ZREG ? MAXD : 1];  // in the stage type: they only feed the stage sums
        if constexpr (ZREG) {
#pragma unroll
            for (int dd = 0; dd < MAXD; ++dd) zr[dd] = cast<T>(zval(p, dd));
        }
        auto zk = [&](int dd) -> cx<T> {
            if constexpr (ZREG) return zr[dd];
            else return cast<T>(zval(p, dd));
        };
        cx<T> Hs[4], rs[4], K[4], acc[4];
        auto hstage = [&](int j) {
#pragma unroll
            for (int q = 0; q < 4; ++q) Hs[q] = {(T)0, (T)0};
#pragma unroll
            for (int dd = 0; dd < MAXD; ++dd) {
                const cx<T> zz = zk(dd);
                const cx<T>* gp = G + (size_t)(j * MAXD + dd) * E;
                cx<T> g0, g1, g2, g3;
                if constexpr (MF) {
                    cx<T> gc[4];
                    TT::ld4(gp, 0, bl, gc);
                    g0 = gc[0]; g1 = gc[1]; g2 = gc[2]; g3 = gc[3];
                } else {
                    TT::ld2(gp, i0, j0, g0, g1);
                    TT::ld2(gp, i0 + 1, j0, g2, g3);
                }
                cfma(Hs[0], zz, g0);
                cfma(Hs[1], zz, g1);
                cfma(Hs[2], zz, g2);
                cfma(Hs[3], zz, g3);
            }
        };
        auto deriv = [&]() {
            if constexpr (MF) {
                // K = -i C with C = H X - X H, on v_mfma_f32_4x4x1f32: the 4 lanes of
                // a k-point form one block, lane 4b + i supplies A[i], lane 4b + j B[j],
                // and lane 4b + j receives D[r][j], r = 0..3 (its own column). With
                // H, X Hermitian, C_ij = sum_m conj(H_mi) X_mj - conj(X_mi) H_mj, so
                // both operands are the lanes' own column elements (m = 0..3):
                //   Re C = P - P^T,                    P  = sum_m Hr_i Xr_j + Hi_i Xi_j
                //   Im C = (Q - S) + (Q - S)^T,        Q  = sum_m Hr_i Xi_j,  S = sum_m Hi_i Xr_j
                // Each transpose is its own accumulator fed the same products in the
                // same order (P^T = sum_m Xr_i Hr_j + Xi_i Hi_j), so Re C is exactly
                // antisymmetric, Im C exactly symmetric and K exactly Hermitian, as in
                // the M - M^dagger form (32 MFMAs per stage for the wave's 16 k-points).
                f32x4 p = {0, 0, 0, 0}, pt = {0, 0, 0, 0}, qq = {0, 0, 0, 0}, qt = {0, 0, 0, 0},
                      ss = {0, 0, 0, 0}, st = {0, 0, 0, 0};
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const float hr = Hs[m].re, hi = Hs[m].im, xr = rs[m].re, xi = rs[m].im;
                    p = __builtin_amdgcn_mfma_f32_4x4x1f32(hr, xr, p, 0, 0, 0);
                    pt = __builtin_amdgcn_mfma_f32_4x4x1f32(xr, hr, pt, 0, 0, 0);
                    qq = __builtin_amdgcn_mfma_f32_4x4x1f32(hr, xi, qq, 0, 0, 0);
                    qt = __builtin_amdgcn_mfma_f32_4x4x1f32(xi, hr, qt, 0, 0, 0);
                    ss = __builtin_amdgcn_mfma_f32_4x4x1f32(hi, xr, ss, 0, 0, 0);
                    st = __builtin_amdgcn_mfma_f32_4x4x1f32(xr, hi, st, 0, 0, 0);
                    p = __builtin_amdgcn_mfma_f32_4x4x1f32(hi, xi, p, 0, 0, 0);
                    pt = __builtin_amdgcn_mfma_f32_4x4x1f32(xi, hi, pt, 0, 0, 0);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) K[q] = {(qq[q] - ss[q]) + (qt[q] - st[q]), pt[q] - p[q]};
                return;
            } else {
            TT::st2(bufH, i0, j0, Hs[0], Hs[1]);
            TT::st2(bufH, i0 + 1, j0, Hs[2], Hs[3]);
            TT::st2(bufX, i0, j0, rs[0], rs[1]);
            TT::st2(bufX, i0 + 1, j0, rs[2], rs[3]);
            wave_lds_sync();
            cx<T> M[4];
            if constexpr (NBP == 16) {
                // 16 x 16 complex product on the matrix cores, one k-point per
                // wave, as three real 16x16x16 products of four v_mfma_*_16x16x4
                // each (P1 = Hr Xr, P2 = Hi Xi, P3 = (Hr + Hi)(Xr + Xi); Re M =
                // P1 - P2, Im M = P3 - P1 - P2: 12 instead of 16 MFMAs). Lane l
                // supplies A[l%16][kc4 + l/16] = H and B[kc4 + l/16][l%16] = X and
                // receives, r = 0..3, D[4 (l/16) + r][l%16] (f32) or
                // D[4 r + l/16][l%16] (f64).
                const int ai = lane & 15, kq = lane >> 4;
                using acc_t = typename std::conditional<std::is_same<T, float>::value, f32x4, f64x4>::type;
                acc_t p1 = {0, 0, 0, 0}, p2 = {0, 0, 0, 0}, p3 = {0, 0, 0, 0};
                auto mfma = [](T a, T b, acc_t c) -> acc_t {
                    if constexpr (std::is_same<T, float>::value) return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
                    else return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
                };
#pragma unroll
                for (int kc = 0; kc < 4; ++kc) {
                    const int kk = 4 * kc + kq;
                    const cx<T> hc = bufH[TT::off(kk, ai)];  // H[ai][kk] = conj(H[kk][ai])
                    const cx<T> x = bufX[TT::off(kk, ai)];   // X[kk][ai]
                    p1 = mfma(hc.re, x.re, p1);                    // Hr Xr
                    p2 = mfma(-hc.im, x.im, p2);                   // Hi Xi, Hi = -hc.im
                    p3 = mfma(hc.re - hc.im, x.re + x.im, p3);     // (Hr + Hi)(Xr + Xi)
                }
                wave_lds_sync();
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = std::is_same<T, float>::value ? 4 * kq + r : 4 * r + kq;
                    bufH[TT::off(row, ai)] = cx<T>{p1[r] - p2[r], p3[r] - p1[r] - p2[r]};
                }
                wave_lds_sync();
                TT::ld2(bufH, i0, j0, M[0], M[1]);
                TT::ld2(bufH, i0 + 1, j0, M[2], M[3]);
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) M[q] = {(T)0, (T)0};
#pragma unroll
                for (int m = 0; m < NBP; ++m) {
                    cx<T> h0, h1, r0, r1;
                    TT::ld2(bufH, m, i0, h0, h1);  // H[m][i0..i0+1] = conj(H[i0..i0+1][m])
                    TT::ld2(bufX, m, j0, r0, r1);  // X[m][j0..j0+1]
                    cfmac(M[0], h0, r0);
                    cfmac(M[1], h0, r1);
                    cfmac(M[2], h1, r0);
                    cfmac(M[3], h1, r1);
                }
                wave_lds_sync();
                TT::st2(bufH, i0, j0, M[0], M[1]);
                TT::st2(bufH, i0 + 1, j0, M[2], M[3]);
            }
            wave_lds_sync();
            cx<T> t0, t1, u0, u1;
            TT::ld2(bufH, j0, i0, t0, t1);      // M[j0][i0], M[j0][i0+1]
            TT::ld2(bufH, j0 + 1, i0, u0, u1);  // M[j0+1][i0], M[j0+1][i0+1]
            // K_ij = -i (M_ij - conj(M_ji))
            K[0] = {M[0].im + t0.im, t0.re - M[0].re};
            K[1] = {M[1].im + u0.im, u0.re - M[1].re};
            K[2] = {M[2].im + t1.im, t1.re - M[2].re};
            K[3] = {M[3].im + u1.im, u1.re - M[3].re};
            wave_lds_sync();
            }
        };
        hstage(0);  // H(E1)
#pragma unroll
        for (int q = 0; q < 4; ++q) rs[q] = cast<T>(rho[q]);
        deriv();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            acc[q] = K[q];
            rs[q] = {fmaT(half, K[q].re, (T)rho[q].re), fmaT(half, K[q].im, (T)rho[q].im)};
        }
        // H(E3), and the midpoint stages' H(E2) = (H(E1) + H(E3)) / 2 (the
        // stage Hamiltonian is affine in the field and E2 = E1 / 2 + E3 / 2)
        cx<T> H3[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) H3[q] = Hs[q];
        hstage(1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const cx<T> h1 = H3[q];
            H3[q] = Hs[q];
            Hs[q] = {(T)0.5 * (h1.re + Hs[q].re), (T)0.5 * (h1.im + Hs[q].im)};
        }
        deriv();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            acc[q].re = fmaT((T)2, K[q].re, acc[q].re); acc[q].im = fmaT((T)2, K[q].im, acc[q].im);
            rs[q] = {fmaT(half, K[q].re, (T)rho[q].re), fmaT(half, K[q].im, (T)rho[q].im)};
        }
        deriv();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            acc[q].re = fmaT((T)2, K[q].re, acc[q].re); acc[q].im = fmaT((T)2, K[q].im, acc[q].im);
            rs[q] = {fmaT(full, K[q].re, (T)rho[q].re), fmaT(full, K[q].im, (T)rho[q].im)};
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) Hs[q] = H3[q];
        deriv();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            acc[q].re += K[q].re; acc[q].im += K[q].im;
            rho[q].re = fmaT(pref, (double)acc[q].re, rho[q].re);
            rho[q].im = fmaT(pref, (double)acc[q].im, rho[q].im);
        }
        }
        if constexpr (KPW_MAX > 1) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const cx<double> r0 = rhoA[0][q];
#pragma unroll
                for (int r = 0; r + 1 < KPW_MAX; ++r) rhoA[r][q] = rhoA[r + 1][q];
                rhoA[KPW_MAX - 1][q] = r0;
            }
        }
        }
        // current of this step, one pass over GV for all k-points p of the lane
        // group: sum_d sum_ij GV[d][a][ij] P_d,ij with the twiddles applied
        // first, P_d,ij = sum_p tw2[m2_p n2_d] conj(rho_p,ij), so each axis costs
        // one complex MAC per element, not one per element and k-point (8 bands,
        // two k-points per group: 120 x 120 FP32 50.4 -> 47.7 ms, FP64 72.9 ->
        // 69.9 ms). One code path per kernel: keeping the per-k order beside it
        // for groups that carry one k-point at run time made the FP32 stage
        // block slower (53.3 ms); kernels of one k-point per group keep it.
#pragma unroll
        for (int dd = 0; dd < MAXD; ++dd) {
            if (!((wmask >> dd) & 1)) continue;  // no velocity block at this n2 (uniform)
            if constexpr (KPW_MAX == 1) {
                // one k-point per group (4 bands): the twiddle after the sum
                const cx<double> zz = zval(0, dd);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    if (!((d.vmask >> a) & 1)) continue;
                    const cx<double>* vp = GV + (size_t)(wrank(dd) * 3 + a) * E;
                    cx<double> v0, v1, v2, v3;
                    if constexpr (MF) {
                        cx<double> vc[4];
                        TD::ld4(vp, 0, bl, vc);
                        v0 = vc[0]; v1 = vc[1]; v2 = vc[2]; v3 = vc[3];
                    } else {
                        TD::ld2(vp, i0, j0, v0, v1);
                        TD::ld2(vp, i0 + 1, j0, v2, v3);
                    }
                    cx<double> s = {0, 0};
                    cfmac(s, rhoA[0][0], v0);
                    cfmac(s, rhoA[0][1], v1);
                    cfmac(s, rhoA[0][2], v2);
                    cfmac(s, rhoA[0][3], v3);
                    cfma(jv[a], zz, s);
                }
            } else {
                cx<double> P[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
#pragma unroll
                for (int p = 0; p < KPW_MAX; ++p) {
                    if (p < KPW) {
                        const cx<double> zz = zval(p, dd);
#pragma unroll
                        for (int q = 0; q < 4; ++q) cfmac(P[q], rhoA[p][q], zz);
                    }
                }
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    if (!((d.vmask >> a) & 1)) continue;
                    const cx<double>* vp = GV + (size_t)(wrank(dd) * 3 + a) * E;
                    cx<double> v0, v1, v2, v3;
                    TD::ld2(vp, i0, j0, v0, v1);
                    TD::ld2(vp, i0 + 1, j0, v2, v3);
                    cx<double> s = {0, 0};  // a short chain per (d, a)
                    cfma(s, v0, P[0]);
                    cfma(s, v1, P[1]);
                    cfma(s, v2, P[2]);
                    cfma(s, v3, P[3]);
                    jv[a].re += s.re;
                    jv[a].im += s.im;
                }
            }
        }
        double jc[6] = {jv[0].re, jv[0].im, jv[1].re, jv[1].im, jv[2].re, jv[2].im};
#pragma unroll
        for (int c = 0; c < 6; ++c) jc[c] = wave_total(jc[c]);
        if (lane == 0) {
#pragma unroll
            for (int c = 0; c < 6; ++c) jw[(jb * W + wave) * 6 + c] = jc[c];
        }
        // slot ti & 1 was last read by the build of step ti, which finished
        // before the previous barrier
        if (DB && ti + 2 < d.nt) weights(ti + 2);
        __syncthreads();
        if (tid < 6) {
            double sacc = 0.0;
            for (int w = 0; w < W; ++w) sacc += jw[(jb * W + w) * 6 + tid];
            ip.jpart[((size_t)ti * ip.nwaves + tile) * 6 + tid] = sacc;
        }
        if (!DB && ti + 1 < d.nt) {
            build(ti + 1, 0);
            __syncthreads();
        }
    }
#pragma unroll
    for (int p = 0; p < KPW_MAX; ++p) {
        const int m2 = m2_of(p);
        const bool kval = p < KPW && m2 < N2;
        const int k = m1 * N2 + (kval ? m2 : 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int ei = eli(q), ej = elj(q);
            if (kval && ei < nb && ej < nb) {
                ip.rho_out[2 * ((size_t)(ei * nb + ej) * N + k)] = rhoA[p][q].re;
                ip.rho_out[2 * ((size_t)(ei * nb + ej) * N + k) + 1] = rhoA[p][q].im;
            }
        }
    }
}

// --------------------------------------------------------------------------
// Launch wrappers. Single instance: descriptor in kernel arguments, block b
// handles tile b. Batch: tiles[b] = (instance, tile of that instance), the
// descriptors live in HBM (uniform, scalar loads).
template <int NB, int MAXU, typename T>
__global__ __launch_bounds__(64) void k_thread_per_k(const SolveDesc d, const InstPtrs ip) {
    thread_per_k_body<NB, MAXU, T>(d, ip, blockIdx.x);
}
template <int NB, int MAXU, typename T>
__global__ __launch_bounds__(64) void k_thread_per_k_batch(const DevInst* __restrict__ insts,
                                                           const int2* __restrict__ tiles) {
    const int2 t = tiles[blockIdx.x];
    thread_per_k_body<NB, MAXU, T>(insts[t.x].d, insts[t.x].p, t.y);
}
template <int NREP, int NX, int NV, typename T>
__global__ __launch_bounds__(64) void k_pair2(const SolveDesc d, const PairDesc pd, const InstPtrs ip) {
    pair2_body<NREP, NX, NV, T>(d, pd, ip, blockIdx.x);
}
template <int NREP, int NX, int NV, typename T>
__global__ __launch_bounds__(64) void k_pair2_batch(const DevInst* __restrict__ insts, const int2* __restrict__ tiles) {
    const int2 t = tiles[blockIdx.x];
    pair2_body<NREP, NX, NV, T>(insts[t.x].d, insts[t.x].pd, insts[t.x].p, t.y);
}
template <int NREP, int NX, int NV, typename T>
__global__ __launch_bounds__(64) void k_bloch2(const SolveDesc d, const PairDesc pd, const InstPtrs ip) {
    bloch2_body<NREP, NX, NV, T>(d, pd, ip, blockIdx.x);
}
template <int NREP, int NX, int NV, typename T>
__global__ __launch_bounds__(64) void k_bloch2_batch(const DevInst* __restrict__ insts, const int2* __restrict__ tiles) {
    const int2 t = tiles[blockIdx.x];
    bloch2_body<NREP, NX, NV, T>(insts[t.x].d, insts[t.x].pd, insts[t.x].p, t.y);
}
template <int NREP, int NX, int NV, typename T>
__global__ __launch_bounds__(128) void k_bloch2w(const SolveDesc d, const PairDesc pd, const InstPtrs ip) {
    bloch2w_body<NREP, NX, NV, T>(d, pd, ip, blockIdx.x);
}
template <int NREP, int NX, int NV, typename T>
__global__ __launch_bounds__(64) void k_bloch2s(const SolveDesc d, const PairDesc pd, const InstPtrs ip) {
    bloch2s_body<NREP, NX, NV, T>(d, pd, ip, blockIdx.x);
}
template <int NREP, int NX, int NV, typename T>
__global__ __launch_bounds__(64) void k_bloch2s_batch(const DevInst* __restrict__ insts, const int2* __restrict__ tiles) {
    const int2 t = tiles[blockIdx.x];
    bloch2s_body<NREP, NX, NV, T>(insts[t.x].d, insts[t.x].pd, insts[t.x].p, t.y);
}
template <int NREP, int NX, int NV, typename T>
__global__ __launch_bounds__(64) void k_pair2s(const SolveDesc d, const PairDesc pd, const InstPtrs ip) {
    pair2s_body<NREP, NX, NV, T>(d, pd, ip, blockIdx.x);
}
template <int NREP, int NX, int NV, typename T>
__global__ __launch_bounds__(64) void k_pair2s_batch(const DevInst* __restrict__ insts, const int2* __restrict__ tiles) {
    const int2 t = tiles[blockIdx.x];
    pair2s_body<NREP, NX, NV, T>(insts[t.x].d, insts[t.x].pd, insts[t.x].p, t.y);
}
template <int NBP, int MAXD, typename T>
__global__ __launch_bounds__(512) void k_sep(const SolveDesc d, const InstPtrs ip) {
    sep_body<NBP, MAXD, T>(d, ip, blockIdx.x);
}
template <int NBP, int MAXD, typename T>
__global__ __launch_bounds__(512) void k_sep_batch(const DevInst* __restrict__ insts, const int2* __restrict__ tiles) {
    const int2 t = tiles[blockIdx.x];
    sep_body<NBP, MAXD, T>(insts[t.x].d, insts[t.x].p, t.y);
}
// The lean FP32 form of 5-8 bands: at most 12 waves per workgroup, so at
// most 168 VGPRs (3 waves per SIMD) instead of the staged form's 185.
constexpr int SEP_LEAN_MAX = 12;
template <int NBP, int MAXD>
__global__ __launch_bounds__(64 * SEP_LEAN_MAX) void k_sep_lean(const SolveDesc d, const InstPtrs ip) {
    sep_body<NBP, MAXD, float, true>(d, ip, blockIdx.x);
}
template <int NBP, int MAXD>
__global__ __launch_bounds__(64 * SEP_LEAN_MAX) void k_sep_lean_batch(const DevInst* __restrict__ insts,
                                                                       const int2* __restrict__ tiles) {
    const int2 t = tiles[blockIdx.x];
    sep_body<NBP, MAXD, float, true>(insts[t.x].d, insts[t.x].p, t.y);
}

template <int EPL, typename T>
__global__ __launch_bounds__(256) void k_elem_per_lane(const SolveDesc d, const InstPtrs ip) {
    elem_body<EPL, T>(d, ip, blockIdx.x);
}
template <int EPL, typename T>
__global__ __launch_bounds__(256) void k_elem_per_lane_batch(const DevInst* __restrict__ insts,
                                                             const int2* __restrict__ tiles) {
    const int2 t = tiles[blockIdx.x];
    elem_body<EPL, T>(insts[t.x].d, insts[t.x].p, t.y);
}
// 33-64 bands: a 1024-lane workgroup per k-point, 4 elements per lane
template <typename T>
__global__ __launch_bounds__(1024) void k_elem_big(const SolveDesc d, const InstPtrs ip) {
    elem_body<4, T, true>(d, ip, blockIdx.x);
}
template <typename T>
__global__ __launch_bounds__(1024) void k_elem_big_batch(const DevInst* __restrict__ insts, const int2* __restrict__ tiles) {
    const int2 t = tiles[blockIdx.x];
    elem_body<4, T, true>(insts[t.x].d, insts[t.x].p, t.y);
}

// J[ti][omap[c]] = (sum over waves of component c, in a fixed order) / N;
// output slots are re/im of x, y, z (0..5), unused slots are zero.
struct OMap {
    int o[6];
    double cst[6];  // added to component c after the 1/N normalisation
};
__global__ __launch_bounds__(64) void k_reduce_j(const double* __restrict__ jpart, int nwaves, int C, OMap om,
                                                 double N, double* __restrict__ J) {
    const int ti = blockIdx.x;
    const int lane = threadIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int w = lane; w < nwaves; w += 64) {
        const double* p = jpart + ((size_t)ti * nwaves + w) * C;
        for (int c = 0; c < C; ++c) acc[c] += p[c];
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[c] = wave_sum(acc[c]);
    if (lane == 0) {
        double out[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < 6; ++c)
            if (c < C && om.o[c] >= 0) out[om.o[c]] = acc[c] / N + om.cst[c];
#pragma unroll
        for (int c = 0; c < 6; ++c) J[(size_t)ti * 6 + c] = out[c];
    }
}

// Batched reduction: block b reduces step rt[b].y of instance rt[b].x.
__global__ __launch_bounds__(64) void k_reduce_j_batch(const DevInst* __restrict__ insts, const int2* __restrict__ rt) {
    const int2 t = rt[blockIdx.x];
    const DevInst& I = insts[t.x];
    const int lane = threadIdx.x, C = I.C, nwaves = I.p.nwaves;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int w = lane; w < nwaves; w += 64) {
        const double* p = I.p.jpart + ((size_t)t.y * nwaves + w) * C;
        for (int c = 0; c < C; ++c) acc[c] += p[c];
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[c] = wave_sum(acc[c]);
    if (lane == 0) {
        const double N = (double)I.d.N;
        double out[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < 6; ++c)
            if (c < C && I.omap[c] >= 0) out[I.omap[c]] = acc[c] / N + I.jcst[c];
#pragma unroll
        for (int c = 0; c < 6; ++c) I.J[(size_t)t.y * 6 + c] = out[c];
    }
}

// --------------------------------------------------------------------------
// Initial state on the device, one thread per k (FP64): H_k from the support
// blocks, the cyclic complex Jacobi eigensolver of the host setup, occupation
// [(Q^dagger H_k Q)_nn < 0] and rho = Q diag(occ) Q^dagger (the T = 0 state of
// src/rdm.cpp:23-38 in the W gauge). H, A, Q are element-major scratch
// [nb*nb][N] so that consecutive k-points coalesce.
__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

__global__ __launch_bounds__(128) void k_init_state(const SolveDesc d, const double* __restrict__ blk,
                                                    const double* __restrict__ tw1, const double* __restrict__ tw2,
                                                    double2* __restrict__ H, double2* __restrict__ A,
                                                    double2* __restrict__ Q, double* __restrict__ rho_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.N) return;
    const int nb = d.nb, nb2 = nb * nb, N = d.N;
    const int m1 = k / d.N2, m2 = k - m1 * d.N2;
    auto at = [&](double2* X, int i, int j) -> double2& { return X[(size_t)(i * nb + j) * N + k]; };
    for (int e = 0; e < nb2; ++e) H[(size_t)e * N + k] = make_double2(0.0, 0.0);
    for (int s = 0; s < d.ns; ++s) {
        if (!(d.opmask[s] & (1u << OP_H))) continue;
        const int j1 = posmod((long)m1 * d.n1[s], d.N1), j2 = posmod((long)m2 * d.n2[s], d.N2);
        const double2 w = zmul(make_double2(tw1[2 * j1], tw1[2 * j1 + 1]), make_double2(tw2[2 * j2], tw2[2 * j2 + 1]));
        const double* b = blk + 2 * ((size_t)s * NOPS + OP_H) * nb2;
        for (int e = 0; e < nb2; ++e)
            H[(size_t)e * N + k] = zadd(H[(size_t)e * N + k], zmul(make_double2(b[2 * e], b[2 * e + 1]), w));
    }
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) {
            at(A, i, j) = at(H, i, j);
            at(Q, i, j) = make_double2(i == j ? 1.0 : 0.0, 0.0);
        }
    for (int it = 0; it < 60; ++it) {
        double off = 0, on = 0;
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j) {
                const double2 x = at(A, i, j);
                (i == j ? on : off) += x.x * x.x + x.y * x.y;
            }
        if (off <= 1e-32 * on || off == 0.0) break;
        for (int p = 0; p < nb - 1; ++p)
            for (int r = p + 1; r < nb; ++r) {
                const double2 apr = at(A, p, r);
                const double g = hypot(apr.x, apr.y);
                if (g == 0.0) continue;
                const double2 u = make_double2(apr.x / g, apr.y / g);
                const double app = at(A, p, p).x, arr = at(A, r, r).x;
                const double th = 0.5 * (arr - app) / g;
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
                // unitary G acting on columns p, r: [c, s; -s conj(u), c conj(u)]
                const double2 g00 = make_double2(c, 0.0), g01 = make_double2(sn, 0.0);
                const double2 g10 = make_double2(-sn * u.x, sn * u.y), g11 = make_double2(c * u.x, -c * u.y);
                for (int i = 0; i < nb; ++i) {
                    double2 x = at(A, i, p), y = at(A, i, r);
                    at(A, i, p) = zadd(zmul(x, g00), zmul(y, g10));
                    at(A, i, r) = zadd(zmul(x, g01), zmul(y, g11));
                    x = at(Q, i, p); y = at(Q, i, r);
                    at(Q, i, p) = zadd(zmul(x, g00), zmul(y, g10));
                    at(Q, i, r) = zadd(zmul(x, g01), zmul(y, g11));
                }
                for (int j = 0; j < nb; ++j) {
                    const double2 x = at(A, p, j), y = at(A, r, j);
                    at(A, p, j) = zadd(zmul(zconj(g00), x), zmul(zconj(g10), y));
                    at(A, r, j) = zadd(zmul(zconj(g01), x), zmul(zconj(g11), y));
                }
                at(A, p, r) = make_double2(0.0, 0.0);
                at(A, r, p) = make_double2(0.0, 0.0);
                at(A, p, p) = make_double2(at(A, p, p).x, 0.0);
                at(A, r, r) = make_double2(at(A, r, r).x, 0.0);
            }
    }
    // occupation from the diagonal of Q^dagger H Q with the original H; the
    // occupied flags go into the (now unused) diagonal of A
    for (int n = 0; n < nb; ++n) {
        double dn = 0.0;
        for (int i = 0; i < nb; ++i) {
            double2 hq = make_double2(0.0, 0.0);
            for (int j = 0; j < nb; ++j) hq = zadd(hq, zmul(at(H, i, j), at(Q, j, n)));
            dn += zmul(zconj(at(Q, i, n)), hq).x;
        }
        at(A, n, n) = make_double2(dn < 0.0 ? 1.0 : 0.0, 0.0);
    }
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) {
            double2 sum = make_double2(0.0, 0.0);
            for (int n = 0; n < nb; ++n)
                if (at(A, n, n).x != 0.0) sum = zadd(sum, zmul(at(Q, i, n), zconj(at(Q, j, n))));
            rho_out[2 * ((size_t)(i * nb + j) * N + k)] = sum.x;
            rho_out[2 * ((size_t)(i * nb + j) * N + k) + 1] = sum.y;
        }
}

// The same initial state for 3 <= nb <= 32 with A and Q in LDS: one group of
// 16 lanes (nb <= 16) or 32 lanes per k-point, 4 or 2 groups per one-wave
// workgroup (34 KiB of LDS at 16 bands, small enough to fit beside other
// solves' propagation kernels; 66 KiB at 32). Lane
// r of a group owns index r: it forms row r of H_k, applies each rotation's
// column update to row r of A and Q and its row update to column r of A, and
// computes the occupation of band r and row r of rho. Each element goes through the
// operations of k_init_state in the same order (rotation parameters and the
// convergence sums are evaluated in every lane from the same LDS values, so a
// group takes one branch), and H_k is formed again for the occupation instead
// of being kept. Rows are padded to nb + 1 elements so that the 16 lanes of a
// column access start 4 banks apart.
constexpr int INIT_MAXNB = 32;
__host__ __device__ inline int init_group(int nb) { return nb <= 16 ? 16 : 32; }
static size_t init_lds_bytes(int nb) { return (size_t)(64 / init_group(nb)) * 2 * nb * (nb + 1) * sizeof(double2); }

__global__ __launch_bounds__(64) void k_init_state_lds(const SolveDesc d, const double* __restrict__ blk,
                                                        const double* __restrict__ tw1, const double* __restrict__ tw2,
                                                        double* __restrict__ rho_out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nb = d.nb, nb2 = nb * nb, N = d.N, ld = nb + 1;
    const int gs = init_group(nb), grp = threadIdx.x / gs, r = threadIdx.x % gs;
    const int k = blockIdx.x * (64 / gs) + grp;
    const bool kv = k < N;
    const int kk = kv ? k : N - 1;  // groups past the end repeat the last k-point and store nothing
    const int m1 = kk / d.N2, m2 = kk - m1 * d.N2;
    double2* A = reinterpret_cast<double2*>(smem) + (size_t)grp * 2 * nb * ld;
    double2* Q = A + nb * ld;
    const bool own = r < nb;
    // row r of H_k: per element, sum over the supports in order
    auto form_h_row = [&]() {
        if (!own) return;
        for (int j = 0; j < nb; ++j) {
            double2 acc = make_double2(0.0, 0.0);
            const int e = r * nb + j;
            for (int s = 0; s < d.ns; ++s) {
                if (!(d.opmask[s] & (1u << OP_H))) continue;
                const int j1 = posmod((long)m1 * d.n1[s], d.N1), j2 = posmod((long)m2 * d.n2[s], d.N2);
                const double2 w =
                    zmul(make_double2(tw1[2 * j1], tw1[2 * j1 + 1]), make_double2(tw2[2 * j2], tw2[2 * j2 + 1]));
                const double* b = blk + 2 * ((size_t)s * NOPS + OP_H) * nb2;
                acc = zadd(acc, zmul(make_double2(b[2 * e], b[2 * e + 1]), w));
            }
            A[r * ld + j] = acc;
        }
    };
    form_h_row();
    if (own)
        for (int j = 0; j < nb; ++j) Q[r * ld + j] = make_double2(r == j ? 1.0 : 0.0, 0.0);
    wave_lds_sync();
    for (int it = 0; it < 60; ++it) {
        double off = 0, on = 0;
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j) {
                const double2 x = A[i * ld + j];
                (i == j ? on : off) += x.x * x.x + x.y * x.y;
            }
        if (off <= 1e-32 * on || off == 0.0) break;
        for (int p = 0; p < nb - 1; ++p)
            for (int q = p + 1; q < nb; ++q) {
                const double2 apq = A[p * ld + q];
                const double g = hypot(apq.x, apq.y);
                if (g == 0.0) continue;
                const double2 u = make_double2(apq.x / g, apq.y / g);
                const double app = A[p * ld + p].x, aqq = A[q * ld + q].x;
                const double th = 0.5 * (aqq - app) / g;
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
                const double2 g00 = make_double2(c, 0.0), g01 = make_double2(sn, 0.0);
                const double2 g10 = make_double2(-sn * u.x, sn * u.y), g11 = make_double2(c * u.x, -c * u.y);
                wave_lds_sync();  // every lane has read the parameters' inputs
                if (own) {  // columns p, q of row r
                    double2 x = A[r * ld + p], y = A[r * ld + q];
                    A[r * ld + p] = zadd(zmul(x, g00), zmul(y, g10));
                    A[r * ld + q] = zadd(zmul(x, g01), zmul(y, g11));
                    x = Q[r * ld + p];
                    y = Q[r * ld + q];
                    Q[r * ld + p] = zadd(zmul(x, g00), zmul(y, g10));
                    Q[r * ld + q] = zadd(zmul(x, g01), zmul(y, g11));
                }
                wave_lds_sync();
                if (own) {  // rows p, q of column r
                    const double2 x = A[p * ld + r], y = A[q * ld + r];
                    A[p * ld + r] = zadd(zmul(zconj(g00), x), zmul(zconj(g10), y));
                    A[q * ld + r] = zadd(zmul(zconj(g01), x), zmul(zconj(g11), y));
                }
                wave_lds_sync();
                if (r == 0) {
                    A[p * ld + q] = make_double2(0.0, 0.0);
                    A[q * ld + p] = make_double2(0.0, 0.0);
                    A[p * ld + p] = make_double2(A[p * ld + p].x, 0.0);
                    A[q * ld + q] = make_double2(A[q * ld + q].x, 0.0);
                }
                wave_lds_sync();
            }
    }
    // occupation of band r from the diagonal of Q^dagger H Q (H formed again in A)
    wave_lds_sync();
    form_h_row();
    wave_lds_sync();
    bool occ = false;
    if (own) {
        double dn = 0.0;
        for (int i = 0; i < nb; ++i) {
            double2 hq = make_double2(0.0, 0.0);
            for (int j = 0; j < nb; ++j) hq = zadd(hq, zmul(A[i * ld + j], Q[j * ld + r]));
            dn += zmul(zconj(Q[i * ld + r]), hq).x;
        }
        occ = dn < 0.0;
    }
    wave_lds_sync();
    if (own) A[r * ld + r] = make_double2(occ ? 1.0 : 0.0, 0.0);
    wave_lds_sync();
    if (own && kv)
        for (int j = 0; j < nb; ++j) {
            double2 sum = make_double2(0.0, 0.0);
            for (int n = 0; n < nb; ++n)
                if (A[n * ld + n].x != 0.0) sum = zadd(sum, zmul(Q[r * ld + n], zconj(Q[j * ld + n])));
            rho_out[2 * ((size_t)(r * nb + j) * N + k)] = sum.x;
            rho_out[2 * ((size_t)(r * nb + j) * N + k) + 1] = sum.y;
        }
}

static void set_lds_limit(const void* fn, size_t lds);

// --------------------------------------------------------------------------
// Allocation-free runs (dev_run_static): the current partials of the
// propagation kernel live in this file-scope device array instead of a
// per-solve buffer. Each GPU holds its own copy for the life of the process;
// its address on a device comes from hipGetSymbolAddress, taken once per
// device when the first solve is created there (host code never passes the
// array's name, which is only a host-side stand-in). Runs on one GPU share it:
// dev_run_static serialises them with g_static_mu.
constexpr size_t STATIC_SCRATCH_DOUBLES = (size_t)8 << 20;  // 64 MiB
static __device__ __align__(256) double scratch[STATIC_SCRATCH_DOUBLES];
constexpr int STATIC_MAX_DEVICES = 64;
static std::mutex g_static_mu;
static double* g_scratch_ptr[STATIC_MAX_DEVICES] = {};

static void static_scratch_setup(int dev) {
    if (dev < 0 || dev >= STATIC_MAX_DEVICES) throw std::runtime_error("device index beyond the static scratch table");
    std::lock_guard<std::mutex> lk(g_static_mu);
    if (g_scratch_ptr[dev]) return;
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the static scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(p);
}

// The device initial state of 3..INIT_MAXNB bands keeps its matrices in LDS;
// 1-2 bands use the thread-per-k kernel over 3 element-major scratch planes.
static bool init_needs_scratch(int nb) { return !(nb >= 3 && nb <= INIT_MAXNB); }

struct DeviceSolve {
    SolveDesc d;
    double dt = 0;
    double *blk = nullptr, *steptab = nullptr, *tw1 = nullptr, *tw2 = nullptr;
    double *rho = nullptr, *rho_tmp = nullptr, *jpart = nullptr, *J = nullptr;
    PairDesc pd{};
    double *pcoef = nullptr, *scoef = nullptr, *psteps = nullptr;
    double* hJ = nullptr;  // pinned host copy of J for asynchronous completion
    int last_C = 6;
    size_t jpart_bytes = 0;
    int device = 0;         // HIP device current at creation (its stream's device)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evk = nullptr;
    bool use_herm = true;   // paired Hermitian kernel when the blocks allow it
    bool use_sep = true;    // row-separable kernel for Hermitian 3 <= nb <= 16
    bool tuned[2] = {false, false};  // autotuned launch key per precision (FP64, FP32)
    // 2 bands: every k-point has the same trace 2 r0 (the Bloch-vector kernel
    // applies); jr0[v] = r0 (m0 + m1) of the origin window block of v slot v
    bool r0_uniform = false;
    // the solve's launches share the device with other solves (served
    // requests): it takes the launch rules of batch members
    bool shared = false;
    double jr0[3] = {0, 0, 0};
    double jr0s[3][2] = {};  // the same for the streamed Bloch-vector kernel: (Re, Im) of the complex current
    int tkey[2][4] = {};
    float last_kernel_ms = 0.f;
};

// Device memory is stream-ordered (hipMallocAsync / hipFreeAsync on the
// solve's own stream): creating or destroying a solve never synchronises
// other solves that run concurrently.
template <typename V>
static double* upload(const std::vector<V>& h, hipStream_t s) {
    double* p = nullptr;
    HIP_CHECK(hipMallocAsync((void**)&p, sizeof(V) * std::max<size_t>(h.size(), 1), s));
    if (!h.empty()) HIP_CHECK(hipMemcpyAsync(p, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice, s));
    return p;
}

static double* dalloc(size_t bytes, hipStream_t s) {
    double* p = nullptr;
    HIP_CHECK(hipMallocAsync((void**)&p, std::max<size_t>(bytes, 16), s));
    return p;
}

// Streamed Bloch-vector pack (layout in sbe_types.h) from the streamed pack.
static void bloch_pack(const double* sc, const PairDesc& pd, double* b) {
    std::fill(b, b + BS_N, 0.0);
    // origin outputs (a, d, Re b, Im b) -> g; current (m_0..3) -> u
    auto g3 = [](const double* h, double* o) { o[0] = 2 * h[2]; o[1] = -2 * h[3]; o[2] = h[0] - h[1]; };
    auto u3 = [](const double* m, double* o) { o[0] = m[2]; o[1] = -m[3]; o[2] = m[0] - m[1]; };
    g3(sc, b);
    for (int x = 0; x < pd.nx; ++x) g3(sc + 4 + 4 * x, b + 3 + 3 * x);
    for (int v = 0; v < pd.nv; ++v) {
        u3(sc + 16 + 8 * v, b + 12 + 3 * v);
        u3(sc + 16 + 8 * v + 4, b + 21 + 3 * v);
    }
    // pair coefficients: output j, coefficient u of (Re w, Im w) at c[2 j + u]
    auto gp = [](const double* c, double* q) {
        for (int u = 0; u < 2; ++u) { q[u] = 2 * c[4 + u]; q[2 + u] = -2 * c[6 + u]; q[4 + u] = c[u] - c[2 + u]; }
    };
    auto up = [](const double* c, double* q) {
        for (int u = 0; u < 2; ++u) { q[u] = c[4 + u]; q[2 + u] = -c[6 + u]; q[4 + u] = c[u] - c[2 + u]; }
    };
    for (int r = 0; r < pd.nrep; ++r) {
        const double* cr = sc + PS_HEAD + PS_REP * r;
        double* o = b + BS_HEAD + BS_REP * r;
        gp(cr, o);
        for (int x = 0; x < pd.nx; ++x) gp(cr + 8 + 8 * x, o + 6 + 6 * x);
        for (int v = 0; v < pd.nv; ++v) {
            up(cr + 32 + 16 * v, o + 24 + 12 * v);          // Re J
            up(cr + 32 + 16 * v + 8, o + 24 + 12 * v + 6);  // Im J
        }
    }
}

DeviceSolve* dev_create(const HostSolve& hs) {
    DeviceSolve* ds = new DeviceSolve();
    try {
        ds->d = hs.d;
        ds->dt = hs.dt;
        HIP_CHECK(hipGetDevice(&ds->device));
        static_scratch_setup(ds->device);
        HIP_CHECK(hipStreamCreateWithFlags(&ds->stream, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreate(&ds->ev0));
        HIP_CHECK(hipEventCreate(&ds->ev1));
        HIP_CHECK(hipEventCreate(&ds->evk));
        ds->blk = upload(hs.blk, ds->stream);
        ds->steptab = upload(hs.steptab, ds->stream);
        ds->tw1 = upload(hs.tw1, ds->stream);
        ds->tw2 = upload(hs.tw2, ds->stream);
        if (!hs.rho0.empty()) {
            ds->rho = upload(hs.rho0, ds->stream);
        } else {  // initial state on the device
            const SolveDesc& d = hs.d;
            const size_t plane = (size_t)d.nb * d.nb * d.N;
            ds->rho = dalloc(sizeof(double) * 2 * plane, ds->stream);
            if (!init_needs_scratch(d.nb)) {  // A and Q in LDS, a lane per row
                const size_t lds = init_lds_bytes(d.nb);
                const int gpb = 64 / init_group(d.nb);
                set_lds_limit((const void*)k_init_state_lds, lds);
                hipLaunchKernelGGL(k_init_state_lds, dim3((d.N + gpb - 1) / gpb), dim3(64), lds, ds->stream, d,
                                   ds->blk, ds->tw1, ds->tw2, ds->rho);
                HIP_CHECK(hipGetLastError());
            } else {
                double2* scratch = (double2*)dalloc(sizeof(double2) * 3 * plane, ds->stream);
                hipLaunchKernelGGL(k_init_state, dim3((d.N + 127) / 128), dim3(128), 0, ds->stream, d, ds->blk,
                                   ds->tw1, ds->tw2, scratch, scratch + plane, scratch + 2 * plane, ds->rho);
                HIP_CHECK(hipGetLastError());
                HIP_CHECK(hipFreeAsync(scratch, ds->stream));
            }
        }
        ds->pd = hs.pd;
        std::vector<double> sc;  // alive until the stream synchronisation below
        if (hs.pd.nrep > 0) {
            if (!hs.pcoef.empty()) ds->pcoef = upload(hs.pcoef, ds->stream);
            if (hs.scoef.size() != (size_t)SC_N) throw std::runtime_error("pair coefficient table of unexpected size");
            sc.assign(BS_OFF + BS_N + (BS_N + 1) / 2, 0.0);
            std::copy(hs.scoef.begin(), hs.scoef.end(), sc.begin());
            float* f = reinterpret_cast<float*>(sc.data() + SC_N);
            for (int i = 0; i < SC_N; ++i) f[i] = (float)hs.scoef[i];
            bloch_pack(hs.scoef.data(), hs.pd, sc.data() + BS_OFF);
            float* fb = reinterpret_cast<float*>(sc.data() + BS_OFF + BS_N);
            for (int i = 0; i < BS_N; ++i) fb[i] = (float)sc[BS_OFF + i];
            ds->scoef = upload(sc, ds->stream);
            ds->psteps = upload(hs.psteps, ds->stream);
        }
        if (hs.d.nb == 2 && hs.pd.nrep > 0) {
            // r0_k = (rho00 + rho11) / 2 of the initial state
            const size_t N = hs.d.N;
            std::vector<double> pq(4 * N);
            if (!hs.rho0.empty()) {
                std::copy(hs.rho0.begin(), hs.rho0.begin() + 2 * N, pq.begin());
                std::copy(hs.rho0.begin() + 6 * N, hs.rho0.begin() + 8 * N, pq.begin() + 2 * N);
            } else {
                HIP_CHECK(hipMemcpyAsync(pq.data(), ds->rho, sizeof(double) * 2 * N, hipMemcpyDeviceToHost, ds->stream));
                HIP_CHECK(hipMemcpyAsync(pq.data() + 2 * N, ds->rho + 6 * N, sizeof(double) * 2 * N,
                                         hipMemcpyDeviceToHost, ds->stream));
                HIP_CHECK(hipStreamSynchronize(ds->stream));
            }
            double lo = 1e300, hi = -1e300, sum = 0.0;
            for (size_t k = 0; k < N; ++k) {
                const double r0 = 0.5 * (pq[2 * k] + pq[2 * N + 2 * k]);
                lo = std::min(lo, r0);
                hi = std::max(hi, r0);
                sum += r0;
            }
            ds->r0_uniform = hi - lo <= 1e-12;
            const double r0 = sum / (double)N;
            for (int v = 0; v < hs.pd.nv; ++v) {
                if (!hs.pcoef.empty()) ds->jr0[v] = r0 * (hs.pcoef[16 + 4 * v] + hs.pcoef[16 + 4 * v + 1]);
                ds->jr0s[v][0] = r0 * (hs.scoef[16 + 8 * v] + hs.scoef[16 + 8 * v + 1]);
                ds->jr0s[v][1] = r0 * (hs.scoef[16 + 8 * v + 4] + hs.scoef[16 + 8 * v + 5]);
            }
        }
        ds->rho_tmp = dalloc(sizeof(double) * 2 * (size_t)hs.d.nb * hs.d.nb * hs.d.N, ds->stream);
        ds->J = dalloc(sizeof(double) * 6 * (size_t)hs.d.nt, ds->stream);
        HIP_CHECK(hipHostMalloc((void**)&ds->hJ, sizeof(double) * 6 * (size_t)hs.d.nt, hipHostMallocDefault));
        HIP_CHECK(hipStreamSynchronize(ds->stream));
    } catch (...) {
        dev_destroy(ds);
        throw;
    }
    return ds;
}

void dev_destroy(DeviceSolve* ds) {
    if (!ds) return;
    for (double* p : {ds->blk, ds->steptab, ds->tw1, ds->tw2, ds->rho, ds->rho_tmp, ds->jpart, ds->J, ds->pcoef,
                      ds->scoef, ds->psteps})
        if (p && ds->stream) (void)hipFreeAsync(p, ds->stream);
    if (ds->stream) (void)hipStreamSynchronize(ds->stream);
    if (ds->hJ) (void)hipHostFree(ds->hJ);
    if (ds->ev0) (void)hipEventDestroy(ds->ev0);
    if (ds->ev1) (void)hipEventDestroy(ds->ev1);
    if (ds->evk) (void)hipEventDestroy(ds->evk);
    if (ds->stream) (void)hipStreamDestroy(ds->stream);
    delete ds;
}

size_t dev_bytes(const HostSolve& hs) {
    const SolveDesc& d = hs.d;
    // propagation partials: at most 6 components per wave; elem tiles hold 256/G k in 4 waves
    const int E = d.nb * d.nb;
    const int G = lanes_per_k(d.nb);
    const int kpb = G >= 256 ? 1 : 256 / G;  // k-points per workgroup of max(256, G) lanes
    const size_t waves = (d.nb <= 2) ? (size_t)(d.N + 63) / 64
                                     : (size_t)((d.N + kpb - 1) / kpb) * (size_t)(std::max(256, G) / 64);
    // rho and its scratch twin (2 x 2 nb^2 N doubles); a device-side initial
    // state of 1-2 bands holds 3 more nb^2 N complex planes while it runs
    const size_t rho = 2 * (size_t)E * d.N;
    const size_t init = (hs.rho0.empty() && init_needs_scratch(d.nb)) ? 6 * (size_t)E * d.N : 0;
    return sizeof(double) * (hs.blk.size() + hs.steptab.size() + hs.tw1.size() + hs.tw2.size() + 2 * rho + init +
                             hs.pcoef.size() + (hs.scoef.empty() ? 0 : (size_t)(BS_OFF + BS_N + (BS_N + 1) / 2)) + 1 +
                             hs.psteps.size() + 6 * (size_t)d.nt +
                             6 * (size_t)d.nt * waves);
}

// Which kernel and template instance a solve runs with.
enum { KIND_THREAD = 1, KIND_PAIR = 2, KIND_ELEM = 3, KIND_PAIRS = 4, KIND_SEP = 6, KIND_BLOCH = 8, KIND_SEPL = 9,
       KIND_BLOCHW = 10, KIND_BLOCHS = 11 };
static bool is_bloch(int kind) { return kind == KIND_BLOCH || kind == KIND_BLOCHW; }
constexpr size_t LDS_MAX = 160 * 1024;
struct LaunchKey {
    int kind = 0, a = 0, b = 0, c = 0, fp32 = 0;
    bool operator<(const LaunchKey& o) const {
        return std::tie(kind, a, b, c, fp32) < std::tie(o.kind, o.a, o.b, o.c, o.fp32);
    }
};

// Candidate launch keys of a solve; an invalid candidate has kind 0.
static bool reg_pairs_ok(const PairDesc& pd) {
    return pd.vpaired && (pd.nx == 0 ? pd.nrep <= 4 : (pd.nrep <= 3 && pd.nx <= 2 && pd.nv <= 2));
}
static LaunchKey key_pair(const DeviceSolve* ds, int fp32) {  // coefficients in registers
    const PairDesc& pd = ds->pd;
    LaunchKey k;
    k.fp32 = fp32;
    if (!(pd.nrep > 0 && ds->use_herm && reg_pairs_ok(pd))) return k;
    k.kind = KIND_PAIR;
    k.a = pd.nrep <= 3 ? 3 : 4;
    if (pd.nx == 0) { k.b = 0; k.c = pd.nv <= 2 ? 2 : 3; }
    else { k.b = 2; k.c = 2; }
    return k;
}
static LaunchKey key_bloch(const DeviceSolve* ds, int fp32) {  // Bloch-vector form of key_pair
    LaunchKey k = key_pair(ds, fp32);
    if (k.kind && ds->r0_uniform) k.kind = KIND_BLOCH;
    else k.kind = 0;
    return k;
}
static LaunchKey key_pairs(const DeviceSolve* ds, int fp32) {  // streamed coefficients
    const PairDesc& pd = ds->pd;
    LaunchKey k;
    k.fp32 = fp32;
    if (!(pd.nrep > 0 && ds->use_herm)) return k;
    k.kind = KIND_PAIRS;
    k.a = pd.nrep <= 12 ? 12 : 32;
    // rbc slots and current axes as the model and the polarisation need them
    // (an in-plane field leaves two rbc axes live; a 3-D velocity, three axes)
    const bool r12 = k.a == 12;
    if (pd.nx == 0 && pd.nv <= 2) { k.b = 0; k.c = 2; }
    else if (pd.nx <= 2 && pd.nv <= 2) { k.b = 2; k.c = 2; }
    else if (r12 && pd.nx == 0 && pd.nv <= 3) { k.b = 0; k.c = 3; }
    else if (r12 && pd.nx <= 2 && pd.nv <= 3) { k.b = 2; k.c = 3; }
    else { k.b = 3; k.c = 3; }
    return k;
}
// Bloch-vector form of key_pairs (the same occupation at every k; 12-pair instances)
static LaunchKey key_blochs(const DeviceSolve* ds, int fp32) {
    LaunchKey k = key_pairs(ds, fp32);
    if (k.kind && ds->r0_uniform && k.a == 12) k.kind = KIND_BLOCHS;
    else k.kind = 0;
    return k;
}
static LaunchKey key_thread(const DeviceSolve* ds, int fp32) {
    const SolveDesc& d = ds->d;
    LaunchKey k;
    k.fp32 = fp32;
    if (!(d.nb <= 2 && d.ns <= 32)) return k;
    k.kind = KIND_THREAD;
    k.a = d.nb;
    k.b = d.ns <= 8 ? 8 : d.ns <= 16 ? 16 : 32;  // MAXU: every support point must fit
    return k;
}
static int device_cus();
// w <= 0: waves per workgroup by the default rule (equal chunks of a k-row,
// at most 8 waves, reduced until the LDS footprint fits). fill (single
// solves): each k-row is split into at least CUs / N1 chunks, so that the
// workgroups cover the CUs, down to 4 waves (4 bands on 120 x 120: 8 -> 4
// waves, FP32 22.4 -> 20.7 ms, FP64 30.2 -> 26.8 ms).
// lean: the lean FP32 form (k_sep_lean, KIND_SEPL; no LDS staging, at most
// 168 VGPRs) of 5-8 bands with up to 12 waves per workgroup (one workgroup,
// 3 waves per SIMD, per CU). For 3-4 bands it measured slower (120 x 120:
// 23.1 ms staged, 29.4 ms lean with 4 waves), so they keep the staged form.
static LaunchKey key_sep(const DeviceSolve* ds, int fp32, int w, bool lean = false, bool fill = false) {
    const SolveDesc& d = ds->d;
    LaunchKey k;
    k.fp32 = fp32;
    if (!(d.nb >= 3 && d.nb <= 16 && d.hermhx && d.nd > 0 && ds->use_sep)) return k;
    const int nbp = d.nb <= 4 ? 4 : d.nb <= 8 ? 8 : 16, gpw = 64 / ((nbp / 2) * (nbp / 2));
    if (lean && !(fp32 && nbp == 8)) return k;
    const int maxd = d.nd <= 5 ? 5 : 8;
    const size_t nstage = (fp32 && nbp <= 8 && !lean) ? (size_t)d.ns * d.nb * d.nb : 0;
    const int wmax = lean ? SEP_LEAN_MAX : 8;
    if (w <= 0) {
        int chunks = (d.N2 + wmax * gpw - 1) / (wmax * gpw);
        w = (d.N2 + chunks * gpw - 1) / (chunks * gpw);
        if (fill && w > 4 && chunks < device_cus() / d.N1) {
            chunks = device_cus() / d.N1;
            w = std::max(4, (d.N2 + chunks * gpw - 1) / (chunks * gpw));
        }
        while (w > 1 && sep_layout(nbp, maxd, fp32 ? 4 : 8, w, nstage, d.nwd, d.ns).bytes > LDS_MAX) --w;
    }
    if (w < 1 || w > (lean ? SEP_LEAN_MAX : 8) || sep_layout(nbp, maxd, fp32 ? 4 : 8, w, nstage, d.nwd, d.ns).bytes > LDS_MAX)
        return k;
    k.kind = lean ? KIND_SEPL : KIND_SEP;
    k.a = nbp;
    k.b = w;
    k.c = maxd;
    return k;
}

// batch: the solve runs in a batch, where 5-8-band FP32 members take the lean
// form (and carry two k-points per lane group on rows long enough, sep_kpw).
static int sep_kpw(const SolveDesc& d, const LaunchKey& k, bool batch = false);
static LaunchKey classify_default(const DeviceSolve* ds, int precision, bool batch = false) {
    const SolveDesc& d = ds->d;
    const int fp32 = precision != 0;
    LaunchKey k;
    if (!reg_pairs_ok(ds->pd) && (k = key_blochs(ds, fp32)).kind) return k;
    if (!reg_pairs_ok(ds->pd) && (k = key_pairs(ds, fp32)).kind) return k;
    if ((k = key_bloch(ds, fp32)).kind) {
        // a single solve whose two-wave workgroups (64 k-points each) leave a
        // SIMD per wave takes the two-wave form: 30 x 30, 900 k-points, 1.01 ->
        // 0.82 ms; at 200 x 200 (625 groups) the one-wave form is 2 % faster
        if (!batch && (d.N + 63) / 64 <= 2 * device_cus()) k.kind = KIND_BLOCHW;
        return k;
    }
    if ((k = key_pair(ds, fp32)).kind) return k;
    if ((k = key_thread(ds, fp32)).kind) return k;
    if (d.nb > 64) throw std::runtime_error("unsupported band count " + std::to_string(d.nb) + " (at most 64)");
    if (batch && (k = key_sep(ds, fp32, 0, true)).kind) return k;
    if ((k = key_sep(ds, fp32, 0, false, !batch)).kind) {
        // a single solve whose lane groups carry one k-point each takes the
        // lean form when it applies (8 bands, 200 x 200: 191 -> 150 ms)
        LaunchKey kw;
        if (sep_kpw(d, k) == 1 && (kw = key_sep(ds, fp32, 0, true)).kind) return kw;
        return k;
    }
    // large supports: the staged FP32 blocks exceed the LDS, the lean form has none
    if ((k = key_sep(ds, fp32, 0, true)).kind) return k;
    k = LaunchKey();
    k.kind = KIND_ELEM;
    k.fp32 = fp32;
    // elements per lane (lanes_per_k lanes per k-point: <= 64 up to 16 bands, 256 beyond)
    k.a = d.nb <= 8 ? 1 : d.nb <= 16 ? 4 : d.nb * d.nb <= 512 ? 2 : 4;
    k.b = d.nb > 32 ? 1 : 0;  // the 1024-lane form
    return k;
}

// The key a run uses: the autotuned one when the solve has been tuned for
// this precision, else the default rule.
static LaunchKey classify(const DeviceSolve* ds, int precision, bool batch = false) {
    const int p = precision != 0;
    if (ds->tuned[p]) {
        LaunchKey k;
        k.kind = ds->tkey[p][0]; k.a = ds->tkey[p][1]; k.b = ds->tkey[p][2]; k.c = ds->tkey[p][3]; k.fp32 = p;
        if ((batch || ds->shared) && k.kind == KIND_BLOCHW) k.kind = KIND_BLOCH;  // no batched two-wave form
        return k;
    }
    return classify_default(ds, precision, batch || ds->shared);
}

static int comps_of(const LaunchKey& k) {
    if (k.kind == KIND_PAIR || is_bloch(k.kind)) return 3;
    if (k.kind == KIND_PAIRS || k.kind == KIND_BLOCHS) return 2 * k.c;
    return 6;
}

static OMap omap_of(const DeviceSolve* ds, const LaunchKey& k) {
    OMap m;
    const bool pair = k.kind == KIND_PAIR || is_bloch(k.kind);
    const bool streamed = k.kind == KIND_PAIRS || k.kind == KIND_BLOCHS;
    for (int c = 0; c < 6; ++c) {
        m.o[c] = (pair || streamed) ? -1 : c;
        m.cst[c] = 0.0;
    }
    if (pair)
        for (int v = 0; v < ds->pd.nv; ++v) m.o[v] = 2 * ds->pd.vaxis[v];
    if (is_bloch(k.kind))
        for (int v = 0; v < ds->pd.nv; ++v) m.cst[v] = ds->jr0[v];
    if (streamed)
        for (int v = 0; v < ds->pd.nv; ++v) { m.o[2 * v] = 2 * ds->pd.vaxis[v]; m.o[2 * v + 1] = 2 * ds->pd.vaxis[v] + 1; }
    if (k.kind == KIND_BLOCHS)
        for (int v = 0; v < ds->pd.nv; ++v) { m.cst[2 * v] = ds->jr0s[v][0]; m.cst[2 * v + 1] = ds->jr0s[v][1]; }
    return m;
}

static int device_cus() {
    static const int n = [] {
        int dev = 0, c = 0;
        HIP_CHECK(hipGetDevice(&dev));
        HIP_CHECK(hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev));
        return c > 0 ? c : 1;
    }();
    return n;
}

// k-points per lane group of the row-separable kernel (8 and 16 bands; one
// otherwise). Its workgroups (one per CU, by LDS) stay resident for the whole
// time loop, so a solve takes ceil(tiles / CUs) workgroup lifetimes, each
// per step one table build (measured: about 1.2 k-point steps at 16 bands,
// 0.3 at 8, where it overlaps the step) plus KPW k-point steps; the KPW of
// 1, 2, 4 with the least total wins. In a batch the workgroups of many
// instances keep every CU busy and lifetimes do not count: a batched 5-8-band
// solve (staged or lean form) carries two k-points per lane group whenever its
// rows hold two tiles' worth, which halves the table builds and shares the
// current's pass over the velocity tables (64-instance mix 3.23e9 -> 3.69e9
// k-point steps/s). A single solve of the lean form keeps one.
static int sep_kpw(const SolveDesc& d, const LaunchKey& k, bool batch) {
    if ((k.kind != KIND_SEP && k.kind != KIND_SEPL) || sep_kpw_max(k.a) == 1) return 1;
    if (batch && k.a <= 8) return d.N2 >= 2 * k.b * (64 / ((k.a / 2) * (k.a / 2))) ? 2 : 1;
    if (k.kind == KIND_SEPL) return 1;
    const int cus = device_cus();
    const int kpt1 = k.b * (64 / ((k.a / 2) * (k.a / 2)));
    const double build = k.a == 16 ? 1.2 : 0.3;
    int best = 1;
    double bcost = 0.0;
    for (int kp = 1; kp <= sep_kpw_max(k.a); kp *= 2) {
        const long tiles = (long)d.N1 * ((d.N2 + kpt1 * kp - 1) / (kpt1 * kp));
        const double cost = (double)((tiles + cus - 1) / cus) * (build + kp);
        if (kp == 1 || cost < bcost) { best = kp; bcost = cost; }
    }
    return best;
}

// Tiles (workgroups) of one instance and waves per tile.
static int tiles_of(const SolveDesc& d, const LaunchKey& k, int& waves_per_tile, int& block, bool batch = false) {
    if (k.kind == KIND_SEP || k.kind == KIND_SEPL) {
        block = 64 * k.b;
        waves_per_tile = 1;  // one current partial per workgroup and step
        const int kpt = k.b * (64 / ((k.a / 2) * (k.a / 2))) * sep_kpw(d, k, batch);
        return d.N1 * ((d.N2 + kpt - 1) / kpt);
    }
    if (k.kind == KIND_ELEM) {
        const int G = lanes_per_k(d.nb);
        block = std::max(256, G);
        waves_per_tile = block / 64;
        const int kpb = block / G;
        return (d.N + kpb - 1) / kpb;
    }
    if (k.kind == KIND_BLOCHW) {
        block = 128;
        waves_per_tile = 1;  // one current partial per group of 64 k-points and step
        return (d.N + 63) / 64;
    }
    block = 64;
    waves_per_tile = 1;
    return (d.N + 63) / 64;
}

static size_t lds_of(const SolveDesc& d, const LaunchKey& k) {
    if (k.kind == KIND_SEP)
        return sep_layout(k.a, k.c, k.fp32 ? 4 : 8, k.b, (k.fp32 && k.a <= 8) ? (size_t)d.ns * d.nb * d.nb : 0, d.nwd,
                          d.ns)
            .bytes;
    if (k.kind == KIND_SEPL) return sep_layout(k.a, k.c, 4, k.b, 0, d.nwd, d.ns).bytes;
    if (k.kind != KIND_ELEM) return 0;
    const int G = lanes_per_k(d.nb);
    const int kpb = std::max(256, G) / G;
    const size_t esz = k.fp32 ? sizeof(cx<float>) : sizeof(cx<double>);
    return sizeof(cx<double>) * (size_t)kpb * d.ns + esz * (size_t)kpb * 2 * d.nb * d.nb;
}

// Dispatch over the compiled instances. F(kernel_single, kernel_batch) receives
// the two kernel function pointers of the matching instance.
template <typename F>
static void dispatch(const LaunchKey& k, F&& f) {
    auto go = [&](auto single64, auto batch64, auto single32, auto batch32) {
        if (k.fp32) f(single32, batch32); else f(single64, batch64);
    };
    switch (k.kind) {
        case KIND_PAIR:
#define PAIR(R, X, V)                                                                                    \
    if (k.a == R && k.b == X && k.c == V)                                                                \
        return go(k_pair2<R, X, V, double>, k_pair2_batch<R, X, V, double>, k_pair2<R, X, V, float>,     \
                  k_pair2_batch<R, X, V, float>);
            PAIR(3, 0, 2) PAIR(4, 0, 2) PAIR(3, 0, 3) PAIR(4, 0, 3) PAIR(3, 2, 2)
#undef PAIR
            break;
        case KIND_BLOCH:
#define BLOCH(R, X, V)                                                                                   \
    if (k.a == R && k.b == X && k.c == V)                                                                \
        return go(k_bloch2<R, X, V, double>, k_bloch2_batch<R, X, V, double>, k_bloch2<R, X, V, float>,  \
                  k_bloch2_batch<R, X, V, float>);
            BLOCH(3, 0, 2) BLOCH(4, 0, 2) BLOCH(3, 0, 3) BLOCH(4, 0, 3) BLOCH(3, 2, 2)
#undef BLOCH
            break;
        case KIND_BLOCHW:  // single solves only (batches take KIND_BLOCH)
#define BLOCHW(R, X, V)                                                                                  \
    if (k.a == R && k.b == X && k.c == V)                                                                \
        return go(k_bloch2w<R, X, V, double>, k_bloch2_batch<R, X, V, double>, k_bloch2w<R, X, V, float>, \
                  k_bloch2_batch<R, X, V, float>);
            BLOCHW(3, 0, 2) BLOCHW(4, 0, 2) BLOCHW(3, 0, 3) BLOCHW(4, 0, 3) BLOCHW(3, 2, 2)
#undef BLOCHW
            break;
        case KIND_BLOCHS:
#define BLOCHS(R, X, V)                                                                                  \
    if (k.a == R && k.b == X && k.c == V)                                                                \
        return go(k_bloch2s<R, X, V, double>, k_bloch2s_batch<R, X, V, double>, k_bloch2s<R, X, V, float>, \
                  k_bloch2s_batch<R, X, V, float>);
            BLOCHS(12, 0, 2) BLOCHS(12, 2, 2) BLOCHS(12, 0, 3) BLOCHS(12, 2, 3) BLOCHS(12, 3, 3)
#undef BLOCHS
            break;
        case KIND_PAIRS:
#define PAIRS(R, X, V)                                                                                   \
    if (k.a == R && k.b == X && k.c == V)                                                                \
        return go(k_pair2s<R, X, V, double>, k_pair2s_batch<R, X, V, double>, k_pair2s<R, X, V, float>,  \
                  k_pair2s_batch<R, X, V, float>);
            PAIRS(12, 0, 2) PAIRS(12, 2, 2) PAIRS(12, 0, 3) PAIRS(12, 2, 3) PAIRS(12, 3, 3)
            PAIRS(32, 0, 2) PAIRS(32, 2, 2) PAIRS(32, 3, 3)
#undef PAIRS
            break;
        case KIND_THREAD:
#define THR(NB, U)                                                                                           \
    if (k.a == NB && k.b == U)                                                                               \
        return go(k_thread_per_k<NB, U, double>, k_thread_per_k_batch<NB, U, double>, k_thread_per_k<NB, U, float>, \
                  k_thread_per_k_batch<NB, U, float>);
            THR(1, 8) THR(1, 16) THR(1, 32) THR(2, 8) THR(2, 16) THR(2, 32)
#undef THR
            break;
        case KIND_SEP:
#define SEP(NBP, D)                                                                                     \
    if (k.a == NBP && k.c == D)                                                                          \
        return go(k_sep<NBP, D, double>, k_sep_batch<NBP, D, double>, k_sep<NBP, D, float>,              \
                  k_sep_batch<NBP, D, float>);
            SEP(4, 5) SEP(4, 8) SEP(8, 5) SEP(8, 8) SEP(16, 5) SEP(16, 8)
#undef SEP
            break;
        case KIND_SEPL:  // FP32 only
            if (!k.fp32) break;
            if (k.a == 8 && k.c == 5) return f(k_sep_lean<8, 5>, k_sep_lean_batch<8, 5>);
            if (k.a == 8 && k.c == 8) return f(k_sep_lean<8, 8>, k_sep_lean_batch<8, 8>);
            break;
        case KIND_ELEM:
            if (k.b == 1) {
                if (k.a != 4) break;
                return go(k_elem_big<double>, k_elem_big_batch<double>, k_elem_big<float>, k_elem_big_batch<float>);
            }
            if (k.a == 1)
                return go(k_elem_per_lane<1, double>, k_elem_per_lane_batch<1, double>, k_elem_per_lane<1, float>,
                          k_elem_per_lane_batch<1, float>);
            if (k.a == 2)
                return go(k_elem_per_lane<2, double>, k_elem_per_lane_batch<2, double>, k_elem_per_lane<2, float>,
                          k_elem_per_lane_batch<2, float>);
            if (k.a == 4)
                return go(k_elem_per_lane<4, double>, k_elem_per_lane_batch<4, double>, k_elem_per_lane<4, float>,
                          k_elem_per_lane_batch<4, float>);
            break;
    }
    throw std::runtime_error("no compiled kernel for this launch key");
}

// Grows the current-partial buffer on the solve's stream; returns true when
// it was (re)allocated.
static bool ensure_jpart(DeviceSolve* ds, int nwaves, int C) {
    size_t need = sizeof(double) * C * (size_t)ds->d.nt * nwaves;
    if (need <= ds->jpart_bytes) return false;
    if (ds->jpart) HIP_CHECK(hipFreeAsync(ds->jpart, ds->stream));
    ds->jpart = nullptr;
    ds->jpart = dalloc(need, ds->stream);
    ds->jpart_bytes = need;
    return true;
}

static InstPtrs ptrs_of(const DeviceSolve* ds, int nwaves, int kpw) {
    InstPtrs p;
    p.blk = ds->blk; p.steptab = ds->steptab; p.tw1 = ds->tw1; p.tw2 = ds->tw2;
    p.pcoef = ds->pcoef; p.scoef = ds->scoef; p.psteps = ds->psteps; p.rho_in = ds->rho;
    p.rho_out = ds->rho_tmp; p.jpart = ds->jpart; p.nwaves = nwaves; p.kpw = kpw; p.dt = ds->dt;
    return p;
}

static void set_lds_limit(const void* fn, size_t lds) {
    if (lds > LDS_MAX) throw std::runtime_error("LDS footprint exceeds 160 KiB");
    if (lds > 64 * 1024) HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
}

// Launch the propagation kernel of one solve; returns the number of current
// components per wave partial.
// jpart: when given, the current partials go there (the caller checked its
// size) instead of the solve's own buffer.
static int launch_with(DeviceSolve* ds, const LaunchKey& k, const SolveDesc& d, int& nwaves,
                       double* jpart = nullptr) {
    if (k.kind == KIND_THREAD && d.ns > k.b) throw std::runtime_error("thread-per-k kernel: support exceeds MAXU");
    int wpt = 0, block = 0;
    const int ntiles = tiles_of(d, k, wpt, block, ds->shared);
    nwaves = ntiles * wpt;
    const int C = comps_of(k);
    if (!jpart) ensure_jpart(ds, nwaves, C);
    InstPtrs ip = ptrs_of(ds, nwaves, sep_kpw(d, k, ds->shared));
    if (jpart) ip.jpart = jpart;
    const size_t lds = lds_of(d, k);
    dispatch(k, [&](auto single, auto) {
        set_lds_limit((const void*)single, lds);
        if constexpr (std::is_invocable_v<decltype(single), SolveDesc, PairDesc, InstPtrs>)
            hipLaunchKernelGGL(single, dim3(ntiles), dim3(block), lds, ds->stream, d, ds->pd, ip);
        else
            hipLaunchKernelGGL(single, dim3(ntiles), dim3(block), lds, ds->stream, d, ip);
    });
    HIP_CHECK(hipGetLastError());
    return C;
}

static int launch_single(DeviceSolve* ds, int precision, int& nwaves) {
    return launch_with(ds, classify(ds, precision), ds->d, nwaves);
}

// --------------------------------------------------------------------------
// Autotuner. The candidates of a solve are the launch keys whose kernels
// cover it: for the row-separable kernel every workgroup size of 1..8 waves
// that fits LDS; for nb = 2 Hermitian models the register-pair, streamed-pair
// and thread-per-k kernels. Each candidate propagates the first `steps` steps
// from the current state into the scratch state buffer (the state and the
// recorded current are untouched), timed with events on the solve's stream;
// the fastest wins. Winners are cached per solve shape for the process, so a
// batch or a server tunes each shape once.
static std::mutex g_tune_mu;
static std::map<std::string, LaunchKey> g_tune_cache;
static std::atomic<int> g_autotune{0};  // read by every thread that runs a solve

static std::string shape_of(const DeviceSolve* ds, int precision) {
    const SolveDesc& d = ds->d;
    const PairDesc& pd = ds->pd;
    char buf[256];
    std::snprintf(buf, sizeof buf, "%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d", precision != 0, d.nb, d.N1,
                  d.N2, d.ns, d.nd, d.nwd, d.herm, d.hermhx, d.xmask, d.vmask, pd.nrep, pd.nx, pd.nv, pd.vpaired,
                  (int)ds->use_herm, (int)ds->use_sep);
    return buf;
}

static std::vector<LaunchKey> candidates(const DeviceSolve* ds, int precision) {
    const int fp32 = precision != 0;
    std::vector<LaunchKey> out{classify_default(ds, precision)};
    auto add = [&](const LaunchKey& k) {
        if (!k.kind) return;
        for (auto& o : out)
            if (!(o < k) && !(k < o)) return;
        out.push_back(k);
    };
    if (out[0].kind == KIND_SEP || out[0].kind == KIND_SEPL) {
        for (int w = 1; w <= 8; ++w) add(key_sep(ds, fp32, w));
        for (int w = 1; w <= SEP_LEAN_MAX; ++w) add(key_sep(ds, fp32, w, true));
    }
    if (ds->pd.nrep > 0 && ds->use_herm) {
        LaunchKey kb = key_bloch(ds, fp32);
        add(key_blochs(ds, fp32));
        add(kb);
        if (kb.kind) {
            kb.kind = KIND_BLOCHW;
            add(kb);
        }
        add(key_pair(ds, fp32));
        add(key_pairs(ds, fp32));
        add(key_thread(ds, fp32));
    }
    return out;
}

static void set_tuned(DeviceSolve* ds, int p, const LaunchKey& k) {
    ds->tuned[p] = true;
    ds->tkey[p][0] = k.kind; ds->tkey[p][1] = k.a; ds->tkey[p][2] = k.b; ds->tkey[p][3] = k.c;
}

int dev_autotune(DeviceSolve* ds, int precision, int steps) {
    const int p = precision != 0;
    const std::string shape = shape_of(ds, precision);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(shape);
        if (it != g_tune_cache.end()) {
            set_tuned(ds, p, it->second);
            return it->second.kind;
        }
    }
    const std::vector<LaunchKey> cand = candidates(ds, precision);
    SolveDesc d = ds->d;
    d.nt = std::max(1, std::min(d.nt, steps));
    LaunchKey best = cand[0];
    float best_ms = 0.f;
    for (size_t i = 0; i < cand.size(); ++i) {
        float ms = 0.f;
        for (int rep = 0; rep < 2; ++rep) {  // the first launch loads the code object
            int nw = 0;
            HIP_CHECK(hipEventRecord(ds->ev0, ds->stream));
            launch_with(ds, cand[i], d, nw);
            HIP_CHECK(hipEventRecord(ds->evk, ds->stream));
            HIP_CHECK(hipEventSynchronize(ds->evk));
            HIP_CHECK(hipEventElapsedTime(&ms, ds->ev0, ds->evk));
        }
        if (i == 0 || ms < best_ms) { best_ms = ms; best = cand[i]; }
    }
    set_tuned(ds, p, best);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[shape] = best;
    return best.kind;
}

void set_autotune(int on) { g_autotune = on ? 1 : 0; }
int get_autotune() { return g_autotune; }

void dev_launch_key(DeviceSolve* ds, int precision, int* out5) {
    const LaunchKey k = classify(ds, precision);
    out5[0] = k.kind; out5[1] = k.a; out5[2] = k.b; out5[3] = k.c; out5[4] = k.fp32;
}

int dev_kpoints_per_wave(DeviceSolve* ds, int precision) { return sep_kpw(ds->d, classify(ds, precision), ds->shared); }
void dev_batch_launch_key(DeviceSolve* ds, int precision, int* out6) {
    const LaunchKey k = classify(ds, precision, true);
    out6[0] = k.kind; out6[1] = k.a; out6[2] = k.b; out6[3] = k.c; out6[4] = k.fp32;
    out6[5] = sep_kpw(ds->d, k, true);
}
void dev_set_shared(DeviceSolve* ds, bool shared) { ds->shared = shared; }

void dev_launch(DeviceSolve* ds, int precision) {
    const SolveDesc& d = ds->d;
    if (g_autotune && !ds->tuned[precision != 0]) dev_autotune(ds, precision, 32);
    int nwaves = 0;
    HIP_CHECK(hipEventRecord(ds->ev0, ds->stream));
    const int C = launch_single(ds, precision, nwaves);
    HIP_CHECK(hipEventRecord(ds->evk, ds->stream));
    const OMap om = omap_of(ds, classify(ds, precision));
    hipLaunchKernelGGL(k_reduce_j, dim3(d.nt), dim3(64), 0, ds->stream, ds->jpart, nwaves, C, om, (double)d.N, ds->J);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(ds->hJ, ds->J, sizeof(double) * 6 * (size_t)d.nt, hipMemcpyDeviceToHost, ds->stream));
    HIP_CHECK(hipEventRecord(ds->ev1, ds->stream));
}

bool dev_query(DeviceSolve* ds) {
    const hipError_t e = hipEventQuery(ds->ev1);
    if (e == hipErrorNotReady) return false;
    HIP_CHECK(e);
    return true;
}

double dev_complete(DeviceSolve* ds, double* J_out) {
    HIP_CHECK(hipEventSynchronize(ds->ev1));
    std::swap(ds->rho, ds->rho_tmp);
    if (J_out) std::copy(ds->hJ, ds->hJ + 6 * (size_t)ds->d.nt, J_out);
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, ds->ev0, ds->ev1));
    HIP_CHECK(hipEventElapsedTime(&ds->last_kernel_ms, ds->ev0, ds->evk));
    return ms;
}

const double* dev_host_current(const DeviceSolve* ds) { return ds->hJ; }

double dev_run(DeviceSolve* ds, int precision, double* J_out) {
    dev_launch(ds, precision);
    return dev_complete(ds, J_out);
}

size_t static_scratch_bytes() { return sizeof(double) * STATIC_SCRATCH_DOUBLES; }

// dev_run without allocating: the launch key of the default rule (or the
// tuned one; no tuning here), the current partials in the static scratch of
// the solve's device, the reduced current in the solve's own J buffer, copied
// through its pinned host buffer into J_out. Refuses launches whose partials
// exceed the scratch. Serialised process-wide: the scratch is shared.
double dev_run_static(DeviceSolve* ds, int precision, double* J_out) {
    const SolveDesc& d = ds->d;
    const LaunchKey k = classify(ds, precision);
    int wpt = 0, block = 0;
    const size_t parts = (size_t)tiles_of(d, k, wpt, block, ds->shared) * wpt;
    const size_t need = (size_t)comps_of(k) * d.nt * parts;
    if (need > STATIC_SCRATCH_DOUBLES)
        throw std::runtime_error("allocation-free run: the current partials need " + std::to_string(need * 8) +
                                 " bytes (" + std::to_string(d.nt) + " steps x " + std::to_string(parts) +
                                 " partials x " + std::to_string(comps_of(k)) +
                                 " components), above the static scratch of " +
                                 std::to_string(static_scratch_bytes()) + " bytes");
    std::lock_guard<std::mutex> lk(g_static_mu);
    double* ptr = g_scratch_ptr[ds->device];
    if (!ptr) throw std::runtime_error("allocation-free run: static scratch not set up on this device");
    int nwaves = 0;
    HIP_CHECK(hipEventRecord(ds->ev0, ds->stream));
    const int C = launch_with(ds, k, d, nwaves, ptr);
    HIP_CHECK(hipEventRecord(ds->evk, ds->stream));
    const OMap om = omap_of(ds, k);
    hipLaunchKernelGGL(k_reduce_j, dim3(d.nt), dim3(64), 0, ds->stream, ptr, nwaves, C, om, (double)d.N, ds->J);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(ds->hJ, ds->J, sizeof(double) * 6 * (size_t)d.nt, hipMemcpyDeviceToHost, ds->stream));
    HIP_CHECK(hipEventRecord(ds->ev1, ds->stream));
    return dev_complete(ds, J_out);  // waits for the stream: the scratch is free again on return
}


// --------------------------------------------------------------------------
// Batched runs: every solve of the batch advances nt steps; solves that share
// a launch key go into one launch whose tile table interleaves their
// workgroups, longest solves first. Different keys run on separate streams
// and overlap on the device.
struct Batch {
    std::vector<hipStream_t> streams;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // device tables per launch group (instances, tiles, reduction tiles),
    // kept across runs and grown when a run needs more
    std::vector<void*> scratch;
    std::vector<size_t> scratch_bytes;
    float last_ms = 0.f;
    ~Batch() {
        for (void* p : scratch) (void)hipFree(p);
        for (auto s : streams) (void)hipStreamDestroy(s);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

Batch* batch_create() {
    Batch* b = new Batch();
    try {
        HIP_CHECK(hipEventCreate(&b->e0));
        HIP_CHECK(hipEventCreate(&b->e1));
    } catch (...) {
        delete b;
        throw;
    }
    return b;
}
void batch_destroy(Batch* b) { delete b; }

double batch_run(Batch* b, const std::vector<DeviceSolve*>& items, const std::vector<int>& prec,
                 const std::vector<double*>& J_out) {
    if (items.empty()) return 0.0;
    if (prec.size() != items.size() || J_out.size() != items.size())
        throw std::runtime_error("batch_run: members, precisions and outputs differ in number");
    for (size_t i = 0; i < items.size(); ++i)
        for (size_t j = 0; j < i; ++j)
            if (items[i] == items[j]) throw std::runtime_error("batch_run: a solve appears twice");
    std::map<LaunchKey, std::vector<int>> groups;
    std::vector<LaunchKey> keys(items.size());
    for (size_t i = 0; i < items.size(); ++i)
        if (g_autotune && !items[i]->tuned[prec[i] != 0]) dev_autotune(items[i], prec[i], 32);
    for (size_t i = 0; i < items.size(); ++i) {
        keys[i] = classify(items[i], prec[i], true);
        groups[keys[i]].push_back((int)i);
    }
    while (b->streams.size() < groups.size()) {
        hipStream_t s;
        HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        b->streams.push_back(s);
    }
    // finish any pending single-solve work before the batch reads the states
    for (auto* ds : items) HIP_CHECK(hipStreamSynchronize(ds->stream));
    // table buffers: slot 3 g + {0, 1, 2} of group g; a buffer that is too
    // small is replaced (its previous run has completed: the last run ended
    // with a synchronisation of stream 0, which waited for every group)
    auto devalloc = [&](size_t slot, size_t bytes) {
        bytes = std::max<size_t>(bytes, 16);
        if (b->scratch.size() <= slot) {
            b->scratch.resize(slot + 1, nullptr);
            b->scratch_bytes.resize(slot + 1, 0);
        }
        if (b->scratch_bytes[slot] < bytes) {
            if (b->scratch[slot]) HIP_CHECK(hipFree(b->scratch[slot]));
            b->scratch[slot] = nullptr;
            HIP_CHECK(hipMalloc(&b->scratch[slot], bytes));
            b->scratch_bytes[slot] = bytes;
        }
        return b->scratch[slot];
    };
    HIP_CHECK(hipEventRecord(b->e0, b->streams[0]));
    for (size_t g = 1; g < groups.size(); ++g) HIP_CHECK(hipStreamWaitEvent(b->streams[g], b->e0, 0));
    size_t gi = 0;
    for (auto& [key, idx] : groups) {
        hipStream_t st = b->streams[gi++];
        std::vector<DevInst> insts(idx.size());
        std::vector<int2> tiles, rt;
        size_t lds = 0;
        int block = 0;
        // longest solves (most tiles x steps) first
        std::vector<int> order(idx.size());
        for (size_t j = 0; j < idx.size(); ++j) order[j] = (int)j;
        std::vector<double> cost(idx.size());
        for (size_t j = 0; j < idx.size(); ++j) cost[j] = (double)items[idx[j]]->d.nt;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
        std::vector<int> ntiles(idx.size());
        for (size_t j = 0; j < idx.size(); ++j) {
            DeviceSolve* ds = items[idx[j]];
            int wpt = 0;
            ntiles[j] = tiles_of(ds->d, key, wpt, block, true);
            const int nwaves = ntiles[j] * wpt;
            // a partial buffer (re)allocated on the solve's stream must exist
            // before this group's stream uses it
            if (ensure_jpart(ds, nwaves, comps_of(key))) HIP_CHECK(hipStreamSynchronize(ds->stream));
            DevInst& I = insts[j];
            I.d = ds->d;
            I.pd = ds->pd;
            I.p = ptrs_of(ds, nwaves, sep_kpw(ds->d, key, true));
            I.J = ds->J;
            I.C = comps_of(key);
            const OMap om = omap_of(ds, key);
            for (int c = 0; c < 6; ++c) {
                I.omap[c] = om.o[c];
                I.jcst[c] = om.cst[c];
            }
            lds = std::max(lds, lds_of(ds->d, key));
            for (int t = 0; t < ds->d.nt; ++t) rt.push_back(make_int2((int)j, t));
        }
        // interleave the tiles of equally long solves round-robin, longest first
        std::vector<int> next(idx.size(), 0);
        for (size_t pos = 0; pos < order.size();) {
            size_t end = pos;
            while (end < order.size() && cost[order[end]] == cost[order[pos]]) ++end;
            bool any = true;
            while (any) {
                any = false;
                for (size_t q = pos; q < end; ++q) {
                    const int j = order[q];
                    if (next[j] < ntiles[j]) { tiles.push_back(make_int2(j, next[j]++)); any = true; }
                }
            }
            pos = end;
        }
        const size_t g3 = 3 * (gi - 1);
        DevInst* d_insts = (DevInst*)devalloc(g3, sizeof(DevInst) * insts.size());
        int2* d_tiles = (int2*)devalloc(g3 + 1, sizeof(int2) * tiles.size());
        int2* d_rt = (int2*)devalloc(g3 + 2, sizeof(int2) * rt.size());
        HIP_CHECK(hipMemcpyAsync(d_insts, insts.data(), sizeof(DevInst) * insts.size(), hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(d_tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(d_rt, rt.data(), sizeof(int2) * rt.size(), hipMemcpyHostToDevice, st));
        dispatch(key, [&](auto, auto batch) {
            set_lds_limit((const void*)batch, lds);
            hipLaunchKernelGGL(batch, dim3((unsigned)tiles.size()), dim3(block), lds, st, d_insts, d_tiles);
        });
        HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_reduce_j_batch, dim3((unsigned)rt.size()), dim3(64), 0, st, d_insts, d_rt);
        HIP_CHECK(hipGetLastError());
    }
    for (size_t g = 1; g < groups.size(); ++g) {
        hipEvent_t ev;
        HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        HIP_CHECK(hipEventRecord(ev, b->streams[g]));
        HIP_CHECK(hipStreamWaitEvent(b->streams[0], ev, 0));
        HIP_CHECK(hipEventDestroy(ev));
    }
    HIP_CHECK(hipEventRecord(b->e1, b->streams[0]));
    for (size_t i = 0; i < items.size(); ++i) {
        DeviceSolve* ds = items[i];
        std::swap(ds->rho, ds->rho_tmp);
        HIP_CHECK(hipMemcpyAsync(J_out[i], ds->J, sizeof(double) * 6 * (size_t)ds->d.nt, hipMemcpyDeviceToHost,
                                 b->streams[0]));
    }
    HIP_CHECK(hipStreamSynchronize(b->streams[0]));
    HIP_CHECK(hipEventElapsedTime(&b->last_ms, b->e0, b->e1));
    return b->last_ms;
}

void dev_set_variant(DeviceSolve* ds, int variant) {
    ds->tuned[0] = ds->tuned[1] = false;
    ds->use_herm = (variant == 0);
    ds->use_sep = (variant == 0);
}

int dev_kernel_variant(DeviceSolve* ds, int precision) { return classify(ds, precision).kind; }

// Arithmetic of the launched formulation per k-point and RK4 step, counting
// a fused multiply-add as 2 flops. Hermitian nb = 2 kernels: per-k pair phases
// (6 nrep, streamed form only: the register form folds e^{-i k.R} into its
// coefficients), H and rbc sums (16 nrep each), stage fields (24 nx), four
// commutators (4 x 13), stage inputs (24), RK4 accumulation (15), update (8),
// current (nv (16 nrep + 7)). Generic kernels: 8 (8 nb^3 + 4 ns nb^2 + 16 nv nb^2 + 6 ns / 8), with
// 4 nb^3 for the element kernel on Hermitian models (one product per stage).
double dev_flops_per_kstep(DeviceSolve* ds, int precision) {
    const SolveDesc& d = ds->d;
    const int kind = dev_kernel_variant(ds, precision);
    if (is_bloch(kind)) {
        // per pair 3 Bloch outputs x 2 FMAs for g and each rbc slot, stage
        // fields 18 nx, four cross products (4 x 9), stage inputs 18, RK4
        // accumulation 15, update 6, current nv (12 nrep + 5)
        const PairDesc& pd = ds->pd;
        return 12.0 * pd.nrep * (1 + pd.nx) + 18.0 * pd.nx + 36 + 18 + 15 + 6 + pd.nv * (12.0 * pd.nrep + 5);
    }
    if (kind == KIND_BLOCHS) {
        // per pair the twiddled phase (6), 3 Bloch outputs x 2 FMAs for g and
        // each rbc slot, 6 x 2 FMAs per current axis; stage fields 18 nx, four
        // cross products 36, stage inputs 18, RK4 accumulation 15, update 6,
        // current nv x 2 x 5
        const PairDesc& pd = ds->pd;
        return 6.0 * pd.nrep + 12.0 * pd.nrep * (1 + pd.nx) + 24.0 * pd.nrep * pd.nv + 18.0 * pd.nx + 36 + 18 + 15 +
               6 + 10.0 * pd.nv;
    }
    if (kind == KIND_PAIR || kind == KIND_PAIRS) {
        const PairDesc& pd = ds->pd;
        const double jcost = kind == KIND_PAIR ? pd.nv * (16.0 * pd.nrep + 7) : pd.nv * (32.0 * pd.nrep + 14);
        const double phases = kind == KIND_PAIR ? 0.0 : 6.0 * pd.nrep;
        return phases + 16.0 * pd.nrep * (1 + pd.nx) + 24.0 * pd.nx + 52 + 24 + 15 + 8 + jcost;
    }
    int nv = 0, nx = 0;
    for (int a = 0; a < 3; ++a) { nv += (d.vmask >> a) & 1; nx += (d.xmask >> a) & 1; }
    if (kind == KIND_SEP || kind == KIND_SEPL) {
        // padded size P, table depth D (MAXD): stage sums 2 D P^2 cMAC (the
        // midpoint stages average the two), products 4 P^3 cMAC, the
        // M - M^dagger combination, the average and RK4 axpys ~30 P^2 flops, current per 2x2 lane block and value d 4 cMAC for the
        // twiddled state and 3 x 4 cMAC shared by the KPW k-points of the
        // group (3 (4 + 1) cMAC per k-point for 4 bands); plus the per-row tables (90 flops per support and element)
        // amortised over the k-points of a workgroup
        const LaunchKey k = classify(ds, precision);
        const double P = k.a, P2 = P * P, D = k.c, kpw = sep_kpw(d, k, ds->shared);
        const double kpt = k.b * (64.0 / ((P / 2) * (P / 2))) * kpw;
        const double used = d.N2 / (std::ceil(d.N2 / kpt) * kpt);
        const double table = 90.0 * d.ns * d.nb * d.nb;
        // the 4-band FP32 matrix-core form forms H X - X H: two products per stage
        const double prods = sep_mf(k.a, precision ? 4 : 8) ? 8.0 : 4.0;
        return 8.0 * (2 * D * P2 + prods * P2 * P) + 30.0 * P2 + D * (P2 / 4) * (P == 4 ? 120.0 : 32.0 + 96.0 / kpw) + table / (kpt * used);
    }
    const double nb = d.nb;
    // element kernel: one nb^3 product per stage for Hermitian models, two otherwise
    const double prods = (kind == KIND_ELEM && d.hermhx) ? 4.0 : 8.0;
    return 8.0 * (prods * nb * nb * nb + 4 * d.ns * nb * nb + 16.0 * nv * nb * nb) + 6.0 * d.ns;
}

// Arithmetic per k-point and RK4 step that the model fixes, independent of
// the kernel that runs it (flops; an FMA counts 2). Live rbc axes nx are those
// with a non-zero field component and a non-zero block, live current axes nv
// those with a velocity block in the window.
//  2 bands, Hermitian blocks (Bloch vector s, ds/dt = 2 h x s): per +-R pair
//    3 outputs x 2 FMAs for h and each live rbc axis (12 (1 + nx) nrep),
//    stage fields 18 nx, four cross products 36, stage inputs 18, RK4
//    accumulation 15, update 6, current per axis 12 nrep + 5 for a Hermitian
//    window (twice that for a complex one).
//  3+ bands, Hermitian H and rbc: the Fourier sums of H and of each live rbc
//    axis over the upper triangle with +-R pairs folded, per distinct n2 value
//    of the support in the row-separable form (4 nd nb^2 (1 + nx)), stage
//    fields 6 nx nb^2, per stage K = -i (M - M^dagger) with M = H X as three
//    real products (4 x 6 nb^3), M - M^dagger, stage inputs and RK4 update
//    14 nb^2, current per axis and distinct window n2 value 4 nb^2 + 8 (the
//    Hermitian rho folds each element pair into two real x complex MACs).
//  Otherwise (non-Hermitian blocks): two complex products per stage, direct
//    Fourier sums over the ns supports and the window.
double dev_min_flops_per_kstep(DeviceSolve* ds) {
    const SolveDesc& d = ds->d;
    const PairDesc& pd = ds->pd;
    const double nb = d.nb, nb2 = nb * nb;
    int nx = 0, nv = 0;
    for (int a = 0; a < 3; ++a) { nx += (d.xmask >> a) & 1; nv += (d.vmask >> a) & 1; }
    if (d.nb == 2 && pd.nrep > 0) {
        const double cur = pd.nv * (12.0 * pd.nrep + 5) * (pd.vpaired ? 1 : 2);
        return 12.0 * pd.nrep * (1 + pd.nx) + 18.0 * pd.nx + 36 + 18 + 15 + 6 + cur;
    }
    auto distinct_n2 = [&](uint32_t mask) {
        std::vector<int> v;
        for (int s = 0; s < d.ns; ++s)
            if ((d.opmask[s] & mask) && std::find(v.begin(), v.end(), d.n2[s]) == v.end()) v.push_back(d.n2[s]);
        return (double)v.size();
    };
    const uint32_t hx = 1u | (d.xmask << 1), vm = (uint32_t)d.vmask << 4;
    if (d.nb >= 3 && d.hermhx) {
        const double nd = distinct_n2(hx), nwd = distinct_n2(vm);
        return 4.0 * nd * nb2 * (1 + nx) + 6.0 * nx * nb2 + 24.0 * nb2 * nb + 14.0 * nb2 + nv * nwd * (4.0 * nb2 + 8);
    }
    int nsup = 0, nwin = 0;
    for (int s = 0; s < d.ns; ++s) {
        nsup += (d.opmask[s] & hx) != 0;
        nwin += (d.opmask[s] & vm) != 0;
    }
    return 8.0 * (8 * nb2 * nb + (1 + nx) * nsup * nb2 + nv * nwin * nb2) + 24.0 * nx * nb2 + 16.0 * nb2;
}

double dev_last_kernel_ms(DeviceSolve* ds) { return ds ? ds->last_kernel_ms : 0.0; }

void dev_get_rho(DeviceSolve* ds, double* out) {
    const SolveDesc& d = ds->d;
    HIP_CHECK(hipMemcpyAsync(out, ds->rho, sizeof(double) * 2 * (size_t)d.nb * d.nb * d.N, hipMemcpyDeviceToHost,
                             ds->stream));
    HIP_CHECK(hipStreamSynchronize(ds->stream));
}

}  // namespace hhg
