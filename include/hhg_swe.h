/* hhg_swe.h -- C ABI of libwannier.so, the MI355X k-space SBE propagator.
 *
 * The first block is exactly the reference's C ABI (src/cwrapper.cpp:5-28),
 * symbol for symbol, so the reference's ctypes driver (pySWE/pyswe.py:9-98)
 * loads this library unchanged. `cdouble` is std::complex<double>, which is
 * layout-identical to double[2] and to numpy complex128.
 *
 * The second block is additive: error reporting (the reference throws or
 * calls std::exit across the ABI), precision selection and read-back hooks.
 * Handles are opaque. Every entry point catches its own errors; a failed call
 * leaves a message readable with SWESim_get_error / hhg_last_error.
 */
#ifndef HHG_SWE_H
#define HHG_SWE_H

/* cdouble of src/cwrapper.cpp: std::complex<double> in C++ (the library's
 * own definitions are compiled against these declarations), a layout-
 * identical pair of doubles in C. */
#ifdef __cplusplus
#include <complex>
typedef std::complex<double> hhg_cdouble;
extern "C" {
#else
typedef struct { double re, im; } hhg_cdouble;
#endif

typedef struct SWESim SWESim;
typedef struct Settings Settings;

/* ---- reference ABI (src/cwrapper.cpp) ---------------------------------- */
SWESim* SWESim_new(void);                                   /* cwrapper.cpp:7  */
void SWESim_run_simulation(SWESim* swe);                    /* cwrapper.cpp:8  */
void SWESim_test_files(SWESim* swe);                        /* cwrapper.cpp:9  */
void SWESim_set_path_tb(SWESim* swe, char* path_tb);        /* cwrapper.cpp:10 */
void SWESim_set_settings(SWESim* swe, Settings* settings);  /* cwrapper.cpp:11 */
void SWESim_init(SWESim* swe);                              /* cwrapper.cpp:12 */
void SWESim_restart(SWESim* swe);                           /* cwrapper.cpp:13 */
void SWESim_get_current(SWESim* swe, double* t, hhg_cdouble* jx,
                        hhg_cdouble* jy, hhg_cdouble* jz);  /* cwrapper.cpp:14 */
void SWESim_delete(SWESim* swe);                            /* cwrapper.cpp:15 */

Settings* Settings_new(void);                               /* cwrapper.cpp:17 */
void Settings_set_nr1(Settings* s, int val);                /* cwrapper.cpp:18 */
void Settings_set_nr2(Settings* s, int val);                /* cwrapper.cpp:19 */
void Settings_set_nr3(Settings* s, int val);                /* cwrapper.cpp:20 */
void Settings_set_tmax(Settings* s, double val);            /* cwrapper.cpp:21 */
void Settings_set_dt(Settings* s, double val);              /* cwrapper.cpp:22 */
void Settings_set_intensity(Settings* s, double val);       /* cwrapper.cpp:23 */
void Settings_set_lambda(Settings* s, double val);          /* cwrapper.cpp:24 */
void Settings_set_tmax_field(Settings* s, double val);      /* cwrapper.cpp:25 */
void Settings_set_pol_vec(Settings* s, double* val);        /* cwrapper.cpp:26, val[3] */
void Settings_set_phi_vec(Settings* s, double* val);        /* cwrapper.cpp:27, val[3] */

/* ---- extensions (no counterpart in the reference) ----------------------- */
/* SWESim_init without touching the GPU: parses the model and builds the
 * host tables (fields, support, initial state). run_simulation then fails
 * until SWESim_init is called. */
void SWESim_init_host(SWESim* swe);
/* Frees a Settings block (the reference has no such call and leaks it). */
void Settings_delete(Settings* s);
/* 0 = last call on this handle succeeded; otherwise an error code. */
int SWESim_get_status(SWESim* swe);
/* Message of the last failed call on this handle ("" when none). */
const char* SWESim_get_error(SWESim* swe);
/* Message of the last failed call in this thread, on any handle. */
const char* hhg_last_error(void);
/* Current of one run: SWESim_get_current returns the first nt samples (the
 * first run, src/swe.cpp:235-242); this returns run `run` (0 = first, -1 =
 * the last one). Returns 0 on success. */
int SWESim_get_current_run(SWESim* swe, int run, double* t, hhg_cdouble* jx, hhg_cdouble* jy, hhg_cdouble* jz);
/* Number of time steps of the initialised solve (= (int)(tmax/dt)). */
int SWESim_get_nt(SWESim* swe);
int SWESim_get_num_orb(SWESim* swe);
/* Arithmetic of the RK4 stages: 0 = FP64 (default), 1 = FP32 stages with the
 * band basis, initial state, state update and J accumulation in FP64. */
void SWESim_set_precision(SWESim* swe, int precision);
/* Kernel selection: 0 = automatic (default; the Hermitian-specialised
 * kernels when every H and rbc block satisfies B(-R) = B(R)^dagger),
 * 1 = generic kernels, 2 = the reference's real-space algorithm (rho(R), 2-D
 * FFTs of every stage in LDS-resident radix kernels, one RK4 step replayed
 * as a HIP graph; FP64 only; it keeps its own state, started from the
 * initial state on its first run). With HHG_FFTFORM_EAGER=1 in the
 * environment at that first run the FFT form enqueues each step's kernels
 * directly instead of replaying the graph (same kernels and order; for
 * tools that trace kernels launched one by one). */
void SWESim_set_kernel_variant(SWESim* swe, int variant);
/* Kernel the next run launches: 1 = thread per k, 2 = Hermitian nb=2 with
 * register pairs, 3 = element per lane, 4 = Hermitian nb=2 with streamed
 * pairs, 6 = row-separable (Hermitian nb >= 3), 7 = FFT form (variant 2),
 * 8 = Hermitian nb=2 with register pairs in Bloch-vector form (every k-point
 * has the same occupation), 9 = row-separable in its lean FP32 form (5-8
 * bands, batched or one k-point per lane group: no LDS staging of the blocks,
 * up to 12 waves per workgroup), 10 = the Bloch-vector form of 8 on two waves
 * per 64 k-points (single solves of at most two such groups per CU; batch
 * members and served requests take 8), 11 = Hermitian nb=2 with streamed
 * pairs in Bloch-vector form (every k-point has the same occupation, up to 12
 * pairs, complex window current); 0 before SWESim_init. */
int SWESim_get_kernel(SWESim* swe);
/* Autotuner: times every launch configuration that covers this solve (kernel
 * family, workgroup size) on its first `steps` RK4 steps without changing
 * its state or recorded current, and keeps the fastest for later runs at the
 * current precision. Winners are cached per solve shape for the process.
 * Returns the chosen kernel (as SWESim_get_kernel), -1 on error. */
int SWESim_autotune(SWESim* swe, int steps);
/* Launch configuration of the next run: out5 = {kernel, a, b, c, fp32}
 * (template parameters; for kernel 6: padded band count, waves per
 * workgroup, table depth). Returns 0 on success. */
int SWESim_get_launch(SWESim* swe, int* out5);
/* k-points each lane group of the next run carries through a step under one
 * per-row table build on the row-separable kernel: 1, 2 or 4 at 9-16 bands,
 * 1 or 2 at 5-8 bands (the value with the fewest workgroup lifetimes on this
 * device's CUs); 1 for every other kernel. Returns -1 on error. */
int SWESim_get_kpoints_per_wave(SWESim* swe);
/* The same two queries under the launch rules of a batch member (hhg_batch_run)
 * or served request: out6 = {kernel, a, b, c, fp32, k-points per lane group}
 * (5-8-band members carry 2 on rows of two tiles' worth). No counterpart in
 * the reference. Returns 0 on success. */
int SWESim_get_batch_launch(SWESim* swe, int* out6);
/* Process-wide: when on, every solve (single, batched or served) is tuned on
 * its first run as by SWESim_autotune(swe, 32). Default off. */
void hhg_set_autotune(int on);
int hhg_get_autotune(void);
/* Unnormalised in-place 2-D FFTs on the GPU of nch row-major n1 x n2 complex
 * fields in host memory (sign -1 forward, +1 backward): the LDS-resident
 * mixed-radix transform of the FFT form (sides up to 512 with prime factors
 * 2, 3, 5, 7, 11, 13). The transform of src/fftw_helper.cpp:5-33 without the
 * 1/N of ifft3. Returns 0, or -1 with hhg_last_error set. */
int hhg_fft2(double* data, int nch, int n1, int n2, int sign);
/* The same transform on caller-owned device memory (src/fftw_helper.cpp:5-33
 * fft3/ifft3 on data already resident in HBM). A plan holds the twiddle
 * tables and its own stream. exec transforms the nch fields at device_data in
 * place `repeats` times, after checking that device_data is 16-byte aligned
 * device memory inside one allocation of at least nch*n1*n2 complex128; work
 * the caller queued on other streams must be complete. Returns the device
 * time in milliseconds, or -1 with hhg_last_error set. create returns NULL
 * (hhg_last_error set) for unsupported sizes. */
typedef struct hhg_fft2_plan hhg_fft2_plan;
hhg_fft2_plan* hhg_fft2_plan_create(int nch, int n1, int n2);
double hhg_fft2_plan_exec(hhg_fft2_plan* plan, void* device_data, int sign, int repeats);
void hhg_fft2_plan_destroy(hhg_fft2_plan* plan);
/* Arithmetic of the launched formulation per k-point and RK4 step (flops,
 * a fused multiply-add counted as 2). */
double SWESim_flops_per_kstep(SWESim* swe);
/* Arithmetic per k-point and RK4 step that the model itself fixes,
 * independent of the kernel: Hermitian +-R pairs folded, only the live rbc
 * axes (non-zero field component) and current axes, the 2-band dynamics as a
 * Bloch-vector cross product, 3+ bands as K = -i (M - M^dagger) with M = H X
 * in three real products. The reference for roofline fractions. */
double SWESim_min_flops_per_kstep(SWESim* swe);
/* Field tables exactly as the time loop consumes them: t[nt] (au),
 * E[3][nt], A[3][nt]. Returns 0 on success. */
int SWESim_get_fields(SWESim* swe, double* t, double* E, double* A);
/* Current k-space density matrices rho_k, [nr1*nr2][nb][nb], k = m1*nr2+m2
 * (after SWESim_init_host: the initial state). Returns 0 on success. */
int SWESim_get_rho_k(SWESim* swe, hhg_cdouble* rho);
/* Device time of the last SWESim_run_simulation in milliseconds (HIP
 * events around the propagation kernels, J reduction included). */
double SWESim_last_run_ms(SWESim* swe);
/* Device time of the propagation kernel alone in the last run (ms). */
double SWESim_last_kernel_ms(SWESim* swe);

/* ---- allocation-free run (no counterpart in the reference) --------------
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.
 * SWESim_run_simulation_static propagates nt steps from the current state
 * like SWESim_run_simulation and writes this run's J(t) to the caller's
 * J_out, nt x 6 doubles (re/im of x, y, z per step), instead of appending it
 * to the handle's record. It calls no hipMalloc/hipFree: the current
 * partials of the propagation kernel (steps x partials x components doubles)
 * live in the library's static device array of 64 MiB per GPU (held until
 * the process ends); a solve whose partials exceed it is refused with an
 * error (e.g. the 2-band 200 x 200 grid at 4096 steps needs 61.4 MB, a
 * 240 x 240 one 88 MB). Calls of one process on one GPU share that array and
 * are serialised by the library. Not available for the FFT form (kernel
 * variant 2). Returns 0, or -1/1 with the handle's error set. */
int SWESim_run_simulation_static(SWESim* swe, double* J_out);

/* ---- batches (no counterpart in the reference) -------------------------
 * A batch propagates many initialised solves together: solves that use the
 * same kernel share one launch whose workgroups interleave their k-tiles;
 * different kernels run concurrently. Each member's current is appended as
 * by SWESim_run_simulation. Functions return 0 on success. */
typedef struct hhg_batch hhg_batch;
hhg_batch* hhg_batch_new(void);
void hhg_batch_delete(hhg_batch* b);
int hhg_batch_add(hhg_batch* b, SWESim* swe);
int hhg_batch_size(hhg_batch* b);
int hhg_batch_run(hhg_batch* b);
double hhg_batch_last_ms(hhg_batch* b);
const char* hhg_batch_get_error(hhg_batch* b);

/* ---- solve server (no counterpart in the reference) ----------------------
 * Continuous batching: requests are prepared on host threads, admitted in
 * arrival order when their HBM footprint fits the budget and one of
 * max_inflight slots is free, and run concurrently on separate HIP streams.
 * The Settings block is copied at submit. Request states: 0 queued,
 * 1 preparing, 2 prepared, 3 running, 4 done, 5 failed. */
typedef struct hhg_server hhg_server;
hhg_server* hhg_server_new(double hbm_budget_gb, int max_inflight, int host_threads);
void hhg_server_delete(hhg_server* s);  /* completes all submitted work first */
/* Admission order among prepared requests: 0 = arrival order, 1 = least
 * estimated device work first (default), the estimate divided by
 * 1 + wait / aging_s so that large requests are still admitted. Returns 0,
 * or -1 with hhg_last_error set. */
int hhg_server_set_policy(hhg_server* s, int policy, double aging_s);
long long hhg_server_submit(hhg_server* s, const char* path_tb, Settings* settings, int precision); /* -1: error */
int hhg_server_wait(hhg_server* s, long long id, double timeout_s);  /* state; timeout < 0 waits */
int hhg_server_state(hhg_server* s, long long id);
int hhg_server_nt(hhg_server* s, long long id);
int hhg_server_result(hhg_server* s, long long id, double* t, hhg_cdouble* jx, hhg_cdouble* jy,
                      hhg_cdouble* jz);
/* out7: submit, prepared, admitted, done (steady-clock seconds), device ms, k-point steps, bytes */
int hhg_server_timing(hhg_server* s, long long id, double* out7);
int hhg_server_release(hhg_server* s, long long id);
/* out10: submitted, completed, failed, inflight, queued, peak inflight, k-point steps, bytes in use,
 * peak bytes, reserved */
int hhg_server_stats(hhg_server* s, double* out10);
const char* hhg_server_error(hhg_server* s, long long id);

#ifdef __cplusplus
}
#endif
#endif /* HHG_SWE_H */
