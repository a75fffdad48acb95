// sbe_fftform.hip -- the reference's real-space algorithm on the GPU ("FFT
// form", kernel variant 2): rho(R) on the nr1 x nr2 supercell and, in every
// RK4 stage, 2-D FFTs of p(R) Heff(R) and of the stage rho, the per-k
// commutator, and the inverse FFT back to R (src/solver.cpp:1428-1726,
// src/fftw_helper.cpp:5-79, src/swe.cpp:183-233, src/observable.cpp:16-30).
//
// It exists as the faithful GPU restatement of the reference loop: a
// full-size cross-check of the per-k kernels and the bandwidth-bound
// formulation (every stage streams 3 nb^2 fields of N points through HBM).
//
// Design:
//  - Fields are stored in FFT order (index n mod N along each axis), which is
//    exactly the reference's fftshift/ifftshift pair for even N: the shifts
//    become the storage convention and cost nothing.
//  - 1-D transforms of a tile of TB lines are LDS-resident: the workgroup
//    loads the tile (coalesced along rows, or TB adjacent columns), runs a
//    mixed-radix Stockham FFT (radices 4, 2, 3, 5, 7, 11, 13) in one LDS
//    buffer (each pass holds its butterflies in registers across a barrier),
//    and stores. Twiddles: the e^{-2 pi i j/N}
//    table of the solve (long-double accurate), staged in LDS with the tile.
//  - One RK4 step (phase, 4 x [form inputs, FFT rows+cols of 2 nb^2 fields,
//    commutator, inverse FFT of nb^2 fields], update, window current) is
//    captured once as a HIP graph and replayed nt times; the kernels read the
//    step index from a device counter that the last node advances.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "sbe_host.h"

namespace hhg {
namespace fftform {

#define FF_CHECK(x)                                                                                          \
    do {                                                                                                     \
        hipError_t e_ = (x);                                                                                 \
        if (e_ != hipSuccess)                                                                                \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " __FILE__ ":" + \
                                     std::to_string(__LINE__));                                              \
    } while (0)

// Lines per workgroup of the column pass: 8 adjacent columns read whole
// 128-byte row segments. Rows are contiguous: a row tile holds about 256
// points (one line of 200-256, two of up to 128, four of up to 64), which
// gives many small workgroups without tiny ones.
// Lines of 257-512 points (the reference's default grid is 400 x 400) run in
// tiles of one row or 4 columns.
constexpr int TB_COLS = 8;
constexpr int SHORT = 256;   // line capacity of the tiles of up to 8 lines
constexpr int MAXLEN = 512;  // longest transform (LDS: (4 + 1) x 512 complex = 40 KiB)
static int rows_per_tile(int L) { return L <= 64 ? 4 : L <= 128 ? 2 : 1; }
static int cols_per_tile(int L) { return L <= SHORT ? TB_COLS : 4; }
constexpr int MAXPASS = 12;

struct Plan {
    int L, n;
    int r[MAXPASS];
};

static Plan make_plan(int L) {
    Plan p{};
    p.L = L;
    int m = L;
    for (int r : {4, 2, 3, 5, 7, 11, 13})
        while (m % r == 0) {
            if (p.n == MAXPASS) throw std::runtime_error("FFT length has too many factors");
            p.r[p.n++] = r;
            m /= r;
        }
    if (m != 1 || L > MAXLEN)
        throw std::runtime_error("FFT form: grid side " + std::to_string(L) +
                                 " must factor into 2, 3, 5, 7, 11, 13 and be at most 512");
    return p;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// Threads of a workgroup that transforms TB lines.
__host__ __device__ constexpr int fft_threads(int TB) { return TB >= 8 ? 256 : TB >= 2 ? 128 : 64; }

// One radix-R pass of the Stockham FFT over the TB lines in buf, in place.
// Butterfly j of a line reads inputs j + q M (q < R), twiddles them by
// e^{sign 2 pi i q k/(Ns R)} (k = j mod Ns), forms a size-R DFT and writes
// outputs (j / Ns) Ns R + k + u Ns. Each thread holds the twiddled inputs of
// its butterflies (at most NBF) in registers across a barrier, so one LDS
// buffer suffices. w: the L roots e^{sign 2 pi i m / L}, in LDS. ML: the
// longest line the tile holds (TB x ML points sets the butterflies per thread).
template <int R, int TB, int NT, int ML>
__device__ __forceinline__ void radix_pass(double2* buf, const double2* w, int L, int Ns) {
    constexpr int NBF = (TB * ML / R + NT - 1) / NT;
    const int M = L / R, total = TB * M, step = L / (Ns * R);
    double2 v[NBF][R];
#pragma unroll
    for (int b = 0; b < NBF; ++b) {
        const int t = threadIdx.x + b * NT;
        if (t < total) {
            const int l = t / M, j = t - l * M, k = j % Ns;
            const double2* s = buf + l * L;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const double2 x = s[j + q * M];
                v[b][q] = q == 0 ? x : cmul(x, w[(q * k * step) % L]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NBF; ++b) {
        const int t = threadIdx.x + b * NT;
        if (t < total) {
            const int l = t / M, j = t - l * M, k = j % Ns;
            double2* d = buf + l * L + (j / Ns) * Ns * R + k;
            auto output = [&](int u) {
                double2 a = v[b][0];
#pragma unroll
                for (int q = 1; q < R; ++q) a = cadd(a, cmul(v[b][q], w[((q * u) % R) * M]));
                d[u * Ns] = a;
            };
            if constexpr (R <= 5) {
#pragma unroll
                for (int u = 0; u < R; ++u) output(u);
            } else {  // the large radices stay a loop: registers for the inputs only
#pragma unroll 1
                for (int u = 0; u < R; ++u) output(u);
            }
        }
    }
    __syncthreads();
}

// All radix passes of the plan over the TB lines of at most ML points in buf
// (NT threads).
template <int TB, int NT, int ML = SHORT>
__device__ __forceinline__ void tile_passes(double2* buf, const double2* w, const Plan& plan) {
    const int L = plan.L;
    int Ns = 1;
    for (int p = 0; p < plan.n; ++p) {
        const int R = plan.r[p];
        switch (R) {
            case 2: radix_pass<2, TB, NT, ML>(buf, w, L, Ns); break;
            case 3: radix_pass<3, TB, NT, ML>(buf, w, L, Ns); break;
            case 4: radix_pass<4, TB, NT, ML>(buf, w, L, Ns); break;
            case 5: radix_pass<5, TB, NT, ML>(buf, w, L, Ns); break;
            case 7: radix_pass<7, TB, NT, ML>(buf, w, L, Ns); break;
            case 11: radix_pass<11, TB, NT, ML>(buf, w, L, Ns); break;
            default: radix_pass<13, TB, NT, ML>(buf, w, L, Ns); break;
        }
        Ns *= R;
    }
}

// Transforms of nch fields of N1 x N2 points (row-major, FFT order) along one
// axis (dim 1: rows of length N2, dim 0: columns of length N1), TB lines per
// workgroup. post: 0 store, 1 store -i x / N (the stage derivative),
// 2 store x / N.
// in and out may alias (in-place row pass): the tile is in LDS before any store.
// ML: the longest line of the instance (SHORT, or MAXLEN for 1 row or 4 columns).
template <int TB, int ML = SHORT>
__global__ __launch_bounds__(fft_threads(TB), TB >= 4 ? 2 : 4) void k_fft_lines(const double2* in, double2* out,
                                                               int nch, int N1, int N2, int dim, int sign,
                                                               const double* __restrict__ tw, Plan plan, int post) {
    extern __shared__ double2 lds[];
    const int L = plan.L;
    const int per_ch = dim == 1 ? N1 : N2;
    const int lines = nch * per_ch;
    const int line0 = blockIdx.x * TB;
    double2* buf = lds;             // the TB lines
    double2* w = lds + TB * L;      // roots of unity of this length and direction
    for (int m = threadIdx.x; m < L; m += blockDim.x)
        w[m] = make_double2(tw[2 * m], sign < 0 ? tw[2 * m + 1] : -tw[2 * m + 1]);
    auto addr = [&](int l, int i) -> size_t {
        const int g = line0 + l, c = g / per_ch, r = g - c * per_ch;
        return dim == 1 ? ((size_t)c * N1 + r) * N2 + i : ((size_t)c * N1 + i) * N2 + r;
    };
    for (int t = threadIdx.x; t < TB * L; t += blockDim.x) {
        int l, i;
        if (dim == 1) { l = t / L; i = t - l * L; }
        else { i = t / TB; l = t - i * TB; }
        if (line0 + l < lines) buf[l * L + i] = in[addr(l, i)];
    }
    __syncthreads();
    tile_passes<TB, fft_threads(TB), ML>(buf, w, plan);
    double2* src = buf;
    const double n = (double)N1 * (double)N2;
    for (int t = threadIdx.x; t < TB * L; t += blockDim.x) {
        int l, i;
        if (dim == 1) { l = t / L; i = t - l * L; }
        else { i = t / TB; l = t - i * TB; }
        if (line0 + l >= lines) continue;
        double2 v = src[l * L + i];
        if (post == 1) v = make_double2(v.y / n, -(v.x / n));  // -i (x / N)
        else if (post == 2) v = make_double2(v.x / n, v.y / n);
        out[addr(l, i)] = v;
    }
}

// p(R) = e^{i A(t).R} on the supercell (src/swe.cpp:225-233); R in bohr as
// (n1 a1 + n2 a2)[Angstrom] / 0.5291772 with n = FFT-order index wrapped to
// [-N/2, N/2).
__global__ void k_phase(const int* __restrict__ step, const double* __restrict__ A, int nt, int N1, int N2,
                        double a10, double a11, double a12, double a20, double a21, double a22,
                        double2* __restrict__ p) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N1 * N2) return;
    const int ti = *step;
    const int j1 = j / N2, j2 = j - j1 * N2;
    const int n1 = j1 < N1 / 2 ? j1 : j1 - N1, n2 = j2 < N2 / 2 ? j2 : j2 - N2;
    double r0 = n1 * a10 + n2 * a20, r1 = n1 * a11 + n2 * a21, r2 = n1 * a12 + n2 * a22;
    r0 *= 1.0 / 0.5291772;
    r1 *= 1.0 / 0.5291772;
    r2 *= 1.0 / 0.5291772;
    const double th = A[ti] * r0 + A[nt + ti] * r1 + A[2 * nt + ti] * r2;
    double s, c;
    sincos(th, &s, &c);
    p[j] = make_double2(c, s);
}

// Stage inputs: F[c] = p (H + E_s . rbc) (c < nb^2), F[nb^2 + c] = rho + coef K_prev.
__global__ void k_form(const int* __restrict__ step, const double* __restrict__ steptab, int stride, int js,
                       double coef, int nb2, int N, const double2* __restrict__ p, const double2* __restrict__ Hd,
                       const double2* __restrict__ Xd, const double2* __restrict__ rho,
                       const double2* __restrict__ Kprev, double2* __restrict__ F) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nb2 * N) return;
    const double* st = steptab + (size_t)(*step) * stride + 3 * js;
    const int j = (int)(i % N);
    const size_t plane = (size_t)nb2 * N;
    double2 h = Hd[i];
    const double2 x0 = Xd[i], x1 = Xd[plane + i], x2 = Xd[2 * plane + i];
    h = make_double2(h.x + st[0] * x0.x, h.y + st[0] * x0.y);
    h = make_double2(h.x + st[1] * x1.x, h.y + st[1] * x1.y);
    h = make_double2(h.x + st[2] * x2.x, h.y + st[2] * x2.y);
    F[i] = cmul(p[j], h);
    const double2 r = rho[i];
    if (Kprev) {
        const double2 k = Kprev[i];
        F[plane + i] = make_double2(r.x + coef * k.x, r.y + coef * k.y);
    } else {
        F[plane + i] = r;
    }
}

// C_k = [H_k, X_k] per k (src/solver.cpp:1461-1476), fields element-major.
__global__ void k_comm(int nb, int N, const double2* __restrict__ F, double2* __restrict__ C) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int nb2 = nb * nb;
    const double2* H = F;
    const double2* X = F + (size_t)nb2 * N;
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) {
            double2 cv = make_double2(0.0, 0.0);
            for (int l = 0; l < nb; ++l) {
                const double2 a = cmul(H[(size_t)(i * nb + l) * N + k], X[(size_t)(l * nb + j) * N + k]);
                const double2 b = cmul(X[(size_t)(i * nb + l) * N + k], H[(size_t)(l * nb + j) * N + k]);
                cv = make_double2(cv.x + (a.x - b.x), cv.y + (a.y - b.y));
            }
            C[(size_t)(i * nb + j) * N + k] = cv;
        }
}

// Fused RK4 stage for nb <= 4, part 1: the row pass of the forward FFTs of the
// 2 nb^2 fields, forming each row as it is loaded (k_form: p(R) (Heff) for the
// first nb^2 channels, the stage rho + coef K_prev for the others).
template <int TB>
__global__ __launch_bounds__(fft_threads(TB), TB >= 4 ? 2 : 4) void k_rows_form(
    const int* __restrict__ step, const double* __restrict__ steptab, int stride, int js, double coef, int nb2,
    int N1, int N2, const double2* __restrict__ p, const double2* __restrict__ Hd, const double2* __restrict__ Xd,
    const double2* __restrict__ rho, const double2* __restrict__ Kprev, const double* __restrict__ tw, Plan plan,
    double2* __restrict__ F) {
    extern __shared__ double2 lds[];
    const int L = plan.L;  // = N2
    const int N = N1 * N2;
    const size_t plane = (size_t)nb2 * N;
    const int lines = 2 * nb2 * N1;
    const int line0 = blockIdx.x * TB;
    double2* buf = lds;
    double2* w = lds + TB * L;
    for (int m = threadIdx.x; m < L; m += blockDim.x) w[m] = make_double2(tw[2 * m], tw[2 * m + 1]);
    const double* st = steptab + (size_t)(*step) * stride + 3 * js;
    for (int t = threadIdx.x; t < TB * L; t += blockDim.x) {
        const int l = t / L, i = t - l * L, g = line0 + l;
        if (g >= lines) continue;
        const int c = g / N1, r = g - c * N1;
        const size_t j = (size_t)r * N2 + i;
        double2 v;
        if (c < nb2) {
            const size_t o = (size_t)c * N + j;
            double2 h = Hd[o];
            const double2 x0 = Xd[o], x1 = Xd[plane + o], x2 = Xd[2 * plane + o];
            h = make_double2(h.x + st[0] * x0.x, h.y + st[0] * x0.y);
            h = make_double2(h.x + st[1] * x1.x, h.y + st[1] * x1.y);
            h = make_double2(h.x + st[2] * x2.x, h.y + st[2] * x2.y);
            v = cmul(p[j], h);
        } else {
            const size_t o = (size_t)(c - nb2) * N + j;
            const double2 rr = rho[o];
            if (Kprev) {
                const double2 k = Kprev[o];
                v = make_double2(rr.x + coef * k.x, rr.y + coef * k.y);
            } else {
                v = rr;
            }
        }
        buf[l * L + i] = v;
    }
    __syncthreads();
    tile_passes<TB, fft_threads(TB)>(buf, w, plan);
    for (int t = threadIdx.x; t < TB * L; t += blockDim.x) {
        const int l = t / L, i = t - l * L, g = line0 + l;
        if (g >= lines) continue;
        const int c = g / N1, r = g - c * N1;
        F[(size_t)c * N + (size_t)r * N2 + i] = buf[l * L + i];
    }
}

// Fused RK4 stage, part 2: for one column n2 of every field, in LDS: the
// forward column FFTs of the 2 nb^2 fields, the commutator C_k = [H_k, X_k] at
// each k of the column (k_comm, src/solver.cpp:1461-1476), and the inverse
// column FFTs of the nb^2 fields of C. The inverse row pass follows.
template <int NB>
__global__ __launch_bounds__(256) void k_cols_comm(int N1, int N2, const double* __restrict__ tw, Plan plan,
                                                   const double2* __restrict__ F, double2* __restrict__ C) {
    constexpr int NB2 = NB * NB, NCH = 2 * NB2;
    extern __shared__ double2 lds[];
    const int L = plan.L;  // = N1
    const int N = N1 * N2;
    const int col = blockIdx.x;
    double2* buf = lds;           // NCH columns of length L
    double2* wf = lds + NCH * L;  // e^{-2 pi i m/L}
    double2* wb = wf + L;         // e^{+2 pi i m/L}
    for (int m = threadIdx.x; m < L; m += blockDim.x) {
        wf[m] = make_double2(tw[2 * m], tw[2 * m + 1]);
        wb[m] = make_double2(tw[2 * m], -tw[2 * m + 1]);
    }
    for (int t = threadIdx.x; t < NCH * L; t += blockDim.x) {
        const int c = t / L, r = t - c * L;
        buf[t] = F[(size_t)c * N + (size_t)r * N2 + col];
    }
    __syncthreads();
    tile_passes<NCH, 256>(buf, wf, plan);
    for (int r = threadIdx.x; r < L; r += blockDim.x) {  // one k per thread; C over the H columns
        double2 h[NB2], x[NB2];
#pragma unroll
        for (int e = 0; e < NB2; ++e) {
            h[e] = buf[e * L + r];
            x[e] = buf[(NB2 + e) * L + r];
        }
#pragma unroll
        for (int i = 0; i < NB; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                double2 cv = make_double2(0.0, 0.0);
#pragma unroll
                for (int l = 0; l < NB; ++l) {
                    const double2 a = cmul(h[i * NB + l], x[l * NB + j]);
                    const double2 b = cmul(x[i * NB + l], h[l * NB + j]);
                    cv = make_double2(cv.x + (a.x - b.x), cv.y + (a.y - b.y));
                }
                buf[(i * NB + j) * L + r] = cv;
            }
    }
    __syncthreads();
    tile_passes<NB2, 256>(buf, wb, plan);
    for (int t = threadIdx.x; t < NB2 * L; t += blockDim.x) {
        const int c = t / L, r = t - c * L;
        C[(size_t)c * N + (size_t)r * N2 + col] = buf[t];
    }
}

// rho += (0.1666666666666 dt)(K0 + 2 K1 + 2 K2 + K3); then the observable
// wrap rho *= conj(p), keeping the 4 x 4 window for the current, and the
// unwrap rho *= p (src/swe.cpp:197-217).
__global__ void k_update(int nb2, int N, int N1, int N2, double pref, const double2* __restrict__ p,
                         const double2* __restrict__ K, double2* __restrict__ rho, double2* __restrict__ win) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nb2 * N) return;
    const size_t plane = (size_t)nb2 * N;
    const double2 k0 = K[i], k1 = K[plane + i], k2 = K[2 * plane + i], k3 = K[3 * plane + i];
    double2 s = make_double2(k0.x + 2.0 * k1.x, k0.y + 2.0 * k1.y);
    s = make_double2(s.x + 2.0 * k2.x, s.y + 2.0 * k2.y);
    s = make_double2(s.x + k3.x, s.y + k3.y);
    double2 r = rho[i];
    r = make_double2(r.x + pref * s.x, r.y + pref * s.y);
    const int j = (int)(i % N), c = (int)(i / N);
    const double2 pj = p[j];
    r = cmul(r, make_double2(pj.x, -pj.y));
    const int j1 = j / N2, j2 = j - j1 * N2;
    const int n1 = j1 < N1 / 2 ? j1 : j1 - N1, n2 = j2 < N2 / 2 ? j2 : j2 - N2;
    if (n1 >= -2 && n1 <= 1 && n2 >= -2 && n2 <= 1) win[((n1 + 2) * 4 + (n2 + 2)) * nb2 + c] = r;
    rho[i] = cmul(r, pj);
}

// J_a[t] = sum over the window (n1, n2 in -2..1) and elements of v_a conj(rho~);
// one workgroup; advances the step counter.
__global__ __launch_bounds__(64) void k_current(int* __restrict__ step, int nb2, const double2* __restrict__ V,
                                                const double2* __restrict__ win, double* __restrict__ J) {
    const int ti = *step;
    const int n = 16 * nb2;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int q = threadIdx.x; q < n; q += 64) {
        const double2 r = win[q];
        for (int a = 0; a < 3; ++a) {
            const double2 v = V[(size_t)a * n + q];
            acc[2 * a] += v.x * r.x + v.y * r.y;
            acc[2 * a + 1] += v.y * r.x - v.x * r.y;
        }
    }
    for (int c = 0; c < 6; ++c)
        for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int c = 0; c < 6; ++c) J[(size_t)ti * 6 + c] = acc[c];
        *step = ti + 1;
    }
}

__global__ void k_set(int* step, int v) { *step = v; }

struct Solve {
    int nb = 0, N1 = 0, N2 = 0, N = 0, nt = 0, stride = 0;
    double dt = 0, lat[6] = {};
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    Plan p1{}, p2{};
    int* step = nullptr;
    double *tw1 = nullptr, *tw2 = nullptr, *steptab = nullptr, *A = nullptr, *J = nullptr;
    double2 *ph = nullptr, *Hd = nullptr, *Xd = nullptr, *V = nullptr, *rho = nullptr, *K = nullptr,
            *F = nullptr, *C = nullptr, *win = nullptr;
    float last_ms = 0.f;
    bool eager = false;  // enqueue each step's kernels directly instead of replaying the graph
};

template <typename T>
static T* dev_upload(const std::vector<T>& h, hipStream_t s) {
    T* p = nullptr;
    FF_CHECK(hipMallocAsync((void**)&p, sizeof(T) * std::max<size_t>(h.size(), 1), s));
    if (!h.empty()) FF_CHECK(hipMemcpyAsync(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
    return p;
}
template <typename T>
static T* dev_zero(size_t n, hipStream_t s) {
    std::vector<T> z(n);
    return dev_upload(z, s);
}

// One pass of 1-D transforms over all lines of one axis (dim 1: rows of
// length N2, dim 0: columns of length N1), tb lines per workgroup.
static void lines_pass(Solve* f, const double2* in, double2* out, int nch, int dim, int sign, int post, int tb) {
    const int L = dim == 1 ? f->N2 : f->N1;
    const int lines = nch * (dim == 1 ? f->N1 : f->N2);
    const int blocks = (lines + tb - 1) / tb;
    const int threads = fft_threads(tb);
    const size_t lds = sizeof(double2) * (tb + 1) * L;
    const double* tw = dim == 1 ? f->tw2 : f->tw1;
    const Plan& p = dim == 1 ? f->p2 : f->p1;
    if (L > SHORT) {  // long lines: one row or 4 columns per tile
        if (tb == 1)
            hipLaunchKernelGGL((k_fft_lines<1, MAXLEN>), dim3(blocks), dim3(threads), lds, f->stream, in, out, nch,
                               f->N1, f->N2, dim, sign, tw, p, post);
        else if (tb == 4)
            hipLaunchKernelGGL((k_fft_lines<4, MAXLEN>), dim3(blocks), dim3(threads), lds, f->stream, in, out, nch,
                               f->N1, f->N2, dim, sign, tw, p, post);
        else
            throw std::runtime_error("FFT form: lines over 256 points run one row or 4 columns per tile");
        return;
    }
    switch (tb) {
        case 1: hipLaunchKernelGGL(k_fft_lines<1>, dim3(blocks), dim3(threads), lds, f->stream, in, out, nch, f->N1,
                                   f->N2, dim, sign, tw, p, post); break;
        case 2: hipLaunchKernelGGL(k_fft_lines<2>, dim3(blocks), dim3(threads), lds, f->stream, in, out, nch, f->N1,
                                   f->N2, dim, sign, tw, p, post); break;
        case 4: hipLaunchKernelGGL(k_fft_lines<4>, dim3(blocks), dim3(threads), lds, f->stream, in, out, nch, f->N1,
                                   f->N2, dim, sign, tw, p, post); break;
        default: hipLaunchKernelGGL(k_fft_lines<8>, dim3(blocks), dim3(threads), lds, f->stream, in, out, nch, f->N1,
                                    f->N2, dim, sign, tw, p, post); break;
    }
}

static void fft2(Solve* f, const double2* in, double2* out, int nch, int sign, int post) {
    // rows (length N2) in place, then columns (length N1) with the post-operation
    lines_pass(f, in, (double2*)in, nch, 1, sign, 0, rows_per_tile(f->N2));
    lines_pass(f, in, out, nch, 0, sign, post, cols_per_tile(f->N1));
}

// One RK4 stage in three launches (nb <= 4, a column of all 2 nb^2 fields in
// LDS): forming + forward row FFTs, forward column FFTs + commutator + inverse
// column FFTs, inverse row FFTs with the -i/N of the stage derivative.
static bool fused_stage_ok(int nb, int N1, int N2) {
    return nb <= 4 && N1 <= SHORT && N2 <= SHORT && sizeof(double2) * (size_t)(2 * nb * nb + 2) * N1 <= 160 * 1024;
}

static void fused_stage(Solve* f, int stride, int js, double coef, const double2* Kprev, double2* Kout) {
    const int nb2 = f->nb * f->nb;
    const int tb = rows_per_tile(f->N2);
    const int lines = 2 * nb2 * f->N1;
    const size_t lds_rows = sizeof(double2) * (tb + 1) * f->N2;
    auto rows = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3((lines + tb - 1) / tb), dim3(fft_threads(tb)), lds_rows, f->stream, f->step,
                           f->steptab, stride, js, coef, nb2, f->N1, f->N2, f->ph, f->Hd, f->Xd, f->rho, Kprev,
                           f->tw2, f->p2, f->F);
    };
    switch (tb) {
        case 1: rows(k_rows_form<1>); break;
        case 2: rows(k_rows_form<2>); break;
        default: rows(k_rows_form<4>); break;
    }
    const size_t lds_cols = sizeof(double2) * (size_t)(2 * nb2 + 2) * f->N1;
    auto cols = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(f->N2), dim3(256), lds_cols, f->stream, f->N1, f->N2, f->tw1, f->p1, f->F,
                           f->C);
    };
    switch (f->nb) {
        case 1: cols(k_cols_comm<1>); break;
        case 2: cols(k_cols_comm<2>); break;
        case 3: cols(k_cols_comm<3>); break;
        default: cols(k_cols_comm<4>); break;
    }
    lines_pass(f, f->C, Kout, nb2, 1, +1, 1, rows_per_tile(f->N2));
}

static void destroy(Solve* f) {
    if (!f) return;
    if (f->exec) (void)hipGraphExecDestroy(f->exec);
    if (f->graph) (void)hipGraphDestroy(f->graph);
    for (void* p : {(void*)f->step, (void*)f->tw1, (void*)f->tw2, (void*)f->steptab, (void*)f->A, (void*)f->J,
                    (void*)f->ph, (void*)f->Hd, (void*)f->Xd, (void*)f->V, (void*)f->rho, (void*)f->K, (void*)f->F,
                    (void*)f->C, (void*)f->win})
        if (p && f->stream) (void)hipFreeAsync(p, f->stream);
    if (f->stream) (void)hipStreamSynchronize(f->stream);
    if (f->e0) (void)hipEventDestroy(f->e0);
    if (f->e1) (void)hipEventDestroy(f->e1);
    if (f->stream) (void)hipStreamDestroy(f->stream);
    delete f;
}

// The kernels of one RK4 step on f's stream (captured once as a graph, or
// enqueued per step in eager mode): phase, 4 stages, update with the
// observable wrap, window current (which advances the device step counter).
static void enqueue_step(Solve* f) {
    const int nb2 = f->nb * f->nb, N = f->N;
    const size_t plane = (size_t)nb2 * N;
    const int blk = 256;
    const unsigned gp = (unsigned)((plane + blk - 1) / blk);
    hipLaunchKernelGGL(k_phase, dim3((N + blk - 1) / blk), dim3(blk), 0, f->stream, f->step, f->A, f->nt, f->N1,
                       f->N2, f->lat[0], f->lat[1], f->lat[2], f->lat[3], f->lat[4], f->lat[5], f->ph);
    const double coef[4] = {0.0, 0.5 * f->dt, 0.5 * f->dt, f->dt};
    const int field[4] = {0, 1, 1, 2};
    const bool fused = fused_stage_ok(f->nb, f->N1, f->N2);
    for (int s = 0; s < 4; ++s) {
        const double2* Kprev = s == 0 ? (const double2*)nullptr : f->K + (size_t)(s - 1) * plane;
        if (fused) {
            fused_stage(f, f->stride, field[s], coef[s], Kprev, f->K + (size_t)s * plane);
            continue;
        }
        hipLaunchKernelGGL(k_form, dim3(gp), dim3(blk), 0, f->stream, f->step, f->steptab, f->stride, field[s],
                           coef[s], nb2, N, f->ph, f->Hd, f->Xd, f->rho, Kprev, f->F);
        fft2(f, f->F, f->F, 2 * nb2, -1, 0);
        hipLaunchKernelGGL(k_comm, dim3((N + blk - 1) / blk), dim3(blk), 0, f->stream, f->nb, N, f->F, f->C);
        fft2(f, f->C, f->K + (size_t)s * plane, nb2, +1, 1);
    }
    hipLaunchKernelGGL(k_update, dim3(gp), dim3(blk), 0, f->stream, nb2, N, f->N1, f->N2, 0.1666666666666 * f->dt,
                       f->ph, f->K, f->rho, f->win);
    hipLaunchKernelGGL(k_current, dim3(1), dim3(64), 0, f->stream, f->step, nb2, f->V, f->win, f->J);
}

}  // namespace fftform

using namespace fftform;

FFTFormSolve* fftform_create(const HostSolve& hs) {
    const SolveDesc& d = hs.d;
    Solve* f = new Solve();
    try {
        f->nb = d.nb; f->N1 = d.N1; f->N2 = d.N2; f->N = d.N; f->nt = d.nt; f->stride = d.stride; f->dt = hs.dt;
        for (int i = 0; i < 6; ++i) f->lat[i] = hs.lat[i];
        f->p1 = make_plan(d.N1);
        f->p2 = make_plan(d.N2);
        FF_CHECK(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
        FF_CHECK(hipEventCreate(&f->e0));
        FF_CHECK(hipEventCreate(&f->e1));
        const int nb2 = d.nb * d.nb, N = d.N;
        const size_t plane = (size_t)nb2 * N;
        // dense operator fields in FFT order (index (n1 mod N1) N2 + (n2 mod N2)), window velocity
        std::vector<double2> Hd(plane), Xd(3 * plane), V(3 * 16 * (size_t)nb2);
        for (int s = 0; s < d.ns; ++s) {
            const int j1 = ((d.n1[s] % d.N1) + d.N1) % d.N1, j2 = ((d.n2[s] % d.N2) + d.N2) % d.N2;
            const size_t j = (size_t)j1 * d.N2 + j2;
            for (int e = 0; e < nb2; ++e) {
                auto blk = [&](int op) {
                    const size_t o = 2 * (((size_t)s * NOPS + op) * nb2 + e);
                    return make_double2(hs.blk[o], hs.blk[o + 1]);
                };
                Hd[(size_t)e * N + j] = blk(OP_H);
                for (int a = 0; a < 3; ++a) Xd[a * plane + (size_t)e * N + j] = blk(OP_X + a);
                if (d.n1[s] >= -2 && d.n1[s] <= 1 && d.n2[s] >= -2 && d.n2[s] <= 1)
                    for (int a = 0; a < 3; ++a)
                        V[(size_t)a * 16 * nb2 + ((d.n1[s] + 2) * 4 + (d.n2[s] + 2)) * nb2 + e] = blk(OP_V + a);
            }
        }
        std::vector<double> A(3 * (size_t)d.nt);
        for (int a = 0; a < 3; ++a)
            for (int t = 0; t < d.nt; ++t) A[(size_t)a * d.nt + t] = hs.A[a][t];
        std::vector<double2> rk(plane);
        for (size_t i = 0; i < plane; ++i) rk[i] = make_double2(hs.rho0[2 * i], hs.rho0[2 * i + 1]);
        f->tw1 = dev_upload(hs.tw1, f->stream);
        f->tw2 = dev_upload(hs.tw2, f->stream);
        f->steptab = dev_upload(hs.steptab, f->stream);
        f->A = dev_upload(A, f->stream);
        f->Hd = dev_upload(Hd, f->stream);
        f->Xd = dev_upload(Xd, f->stream);
        f->V = dev_upload(V, f->stream);
        f->C = dev_upload(rk, f->stream);
        f->rho = dev_zero<double2>(plane, f->stream);
        f->K = dev_zero<double2>(4 * plane, f->stream);
        f->F = dev_zero<double2>(2 * plane, f->stream);
        f->win = dev_zero<double2>(16 * (size_t)nb2, f->stream);
        f->ph = dev_zero<double2>(N, f->stream);
        f->J = dev_zero<double>(6 * (size_t)d.nt, f->stream);
        f->step = dev_zero<int>(1, f->stream);
        // initial state rho(R) = ifftshift(IFFT(rho_k)) / N in FFT order
        fft2(f, f->C, f->rho, nb2, +1, 2);
        FF_CHECK(hipGetLastError());
        // one RK4 step as a graph (HHG_FFTFORM_EAGER=1: enqueued directly every step)
        const char* ev = std::getenv("HHG_FFTFORM_EAGER");
        f->eager = ev && ev[0] == '1';
        if (!f->eager) {
            FF_CHECK(hipStreamBeginCapture(f->stream, hipStreamCaptureModeThreadLocal));
            enqueue_step(f);
            FF_CHECK(hipStreamEndCapture(f->stream, &f->graph));
            FF_CHECK(hipGraphInstantiate(&f->exec, f->graph, nullptr, nullptr, 0));
        }
        FF_CHECK(hipStreamSynchronize(f->stream));
    } catch (...) {
        destroy(f);
        throw;
    }
    return reinterpret_cast<FFTFormSolve*>(f);
}

double fftform_run(FFTFormSolve* fs, double* J_out) {
    Solve* f = reinterpret_cast<Solve*>(fs);
    hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, f->stream, f->step, 0);
    FF_CHECK(hipEventRecord(f->e0, f->stream));
    for (int t = 0; t < f->nt; ++t) {
        if (f->eager) {
            enqueue_step(f);
            FF_CHECK(hipGetLastError());
        } else {
            FF_CHECK(hipGraphLaunch(f->exec, f->stream));
        }
    }
    FF_CHECK(hipEventRecord(f->e1, f->stream));
    FF_CHECK(hipMemcpyAsync(J_out, f->J, sizeof(double) * 6 * (size_t)f->nt, hipMemcpyDeviceToHost, f->stream));
    FF_CHECK(hipStreamSynchronize(f->stream));
    FF_CHECK(hipEventElapsedTime(&f->last_ms, f->e0, f->e1));
    return f->last_ms;
}

void fftform_destroy(FFTFormSolve* fs) { destroy(reinterpret_cast<Solve*>(fs)); }

// e^{-2 pi i j / n}, j < n, rounded from long double
static std::vector<double> twiddle_table(int n) {
    std::vector<double> out(2 * (size_t)n);
    for (int j = 0; j < n; ++j) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * j / n;
        out[2 * j] = (double)cosl(a);
        out[2 * j + 1] = (double)sinl(a);
    }
    return out;
}

struct FFT2Plan {
    Solve* f = nullptr;
    int nch = 0;
};

void fft2_plan_destroy(FFT2Plan* p) {
    if (!p) return;
    destroy(p->f);
    delete p;
}

FFT2Plan* fft2_plan_create(int nch, int n1, int n2) {
    if (nch < 1) throw std::runtime_error("fft2: nch must be positive");
    FFT2Plan* p = new FFT2Plan();
    p->nch = nch;
    p->f = new Solve();
    Solve* f = p->f;
    try {
        f->N1 = n1; f->N2 = n2; f->N = n1 * n2;
        f->p1 = make_plan(n1);
        f->p2 = make_plan(n2);
        FF_CHECK(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
        FF_CHECK(hipEventCreate(&f->e0));
        FF_CHECK(hipEventCreate(&f->e1));
        const std::vector<double> t1 = twiddle_table(n1), t2 = twiddle_table(n2);
        f->tw1 = dev_upload(t1, f->stream);
        f->tw2 = dev_upload(t2, f->stream);
        FF_CHECK(hipStreamSynchronize(f->stream));
    } catch (...) {
        fft2_plan_destroy(p);
        throw;
    }
    return p;
}

double fft2_plan_exec(FFT2Plan* p, void* data, int sign, int repeats) {
    if (!p) throw std::runtime_error("fft2: null plan");
    if (!data) throw std::runtime_error("fft2: null data");
    if (repeats < 1) throw std::runtime_error("fft2: repeats must be positive");
    Solve* f = p->f;
    const size_t bytes = sizeof(double2) * (size_t)p->nch * f->N;
    if (reinterpret_cast<uintptr_t>(data) % alignof(double2))
        throw std::runtime_error("fft2: data must be 16-byte aligned complex128");
    // the kernels touch [data, data + bytes): it must lie in one device allocation
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, data) != hipSuccess || attr.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();
        throw std::runtime_error("fft2: data is not device memory");
    }
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    FF_CHECK(hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)data));
    if ((char*)data + bytes > (char*)base + size)
        throw std::runtime_error("fft2: the allocation holds fewer than nch * n1 * n2 complex values");
    double2* d = reinterpret_cast<double2*>(data);
    FF_CHECK(hipEventRecord(f->e0, f->stream));
    for (int r = 0; r < repeats; ++r) fft2(f, d, d, p->nch, sign < 0 ? -1 : 1, 0);
    FF_CHECK(hipGetLastError());
    FF_CHECK(hipEventRecord(f->e1, f->stream));
    FF_CHECK(hipStreamSynchronize(f->stream));
    float ms = 0.f;
    FF_CHECK(hipEventElapsedTime(&ms, f->e0, f->e1));
    return ms;
}

void fft2_host(double* data, int nch, int n1, int n2, int sign) {
    FFT2Plan* p = fft2_plan_create(nch, n1, n2);
    double2* dd = nullptr;
    try {
        const size_t n = (size_t)nch * n1 * n2;
        FF_CHECK(hipMalloc((void**)&dd, sizeof(double2) * n));
        FF_CHECK(hipMemcpy(dd, data, sizeof(double2) * n, hipMemcpyHostToDevice));
        fft2_plan_exec(p, dd, sign, 1);
        FF_CHECK(hipMemcpy(data, dd, sizeof(double2) * n, hipMemcpyDeviceToHost));
    } catch (...) {
        if (dd) (void)hipFree(dd);
        fft2_plan_destroy(p);
        throw;
    }
    FF_CHECK(hipFree(dd));
    fft2_plan_destroy(p);
}

double fftform_last_ms(FFTFormSolve* fs) { return fs ? reinterpret_cast<Solve*>(fs)->last_ms : 0.0; }

}  // namespace hhg
