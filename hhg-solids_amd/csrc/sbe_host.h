// sbe_host.h -- host-side model, setup tables and the device interface.
#pragma once
#include <complex>
#include <string>
#include <vector>

#include "sbe_types.h"

namespace hhg {

using cplx = std::complex<double>;

// Wannier tight-binding model as read from a W90 _tb.dat file
// (src/wannier_tb.cpp:15-113). Blocks are stored the way the reference
// stores them: file line "m n re im" at [n-1][m-1].
struct TBModel {
    std::string comment;
    double lat[3][3];
    int nb = 0, nR = 0;
    std::vector<int> deg;
    std::vector<int> R;              // nR*3
    std::vector<cplx> H;             // nR*nb*nb, eV
    std::vector<cplx> r[3];          // nR*nb*nb, Angstrom
};
TBModel read_tb(const std::string& path);
// Hermitian eigenproblem of the n x n row-major matrix a by cyclic Jacobi
// rotations: a ends (numerically) diagonal, q holds the eigenvectors as columns.
void herm_jacobi(int n, std::vector<cplx>& a, std::vector<cplx>& q);

// Parameters of one solve; same meaning and units as include/settings.h:5-31.
struct SolveParams {
    int nr1 = 400, nr2 = 400, nr3 = 1;
    double tmax = 90.0, dt = 10.98e-3 / 0.5;
    double lambda_nm = 3000, tmax_field = 80.0, intensity = 5e10;
    double s[3] = {0.96, 0.25, 0.0}, phi[3] = {0.0, 0.0, 0.0};
};

// Everything the device needs for one solve, built on the host in FP64.
struct HostSolve {
    SolveDesc d;
    double dt = 0;                   // t[1]-t[0] (au)
    std::vector<double> blk;         // ns*NOPS*nb*nb complex
    std::vector<double> steptab;     // nt*stride
    std::vector<double> tw1, tw2;    // N1, N2 complex
    std::vector<double> rho0;        // nb*nb*N complex, element-major
    std::vector<double> t, E[3], A[3];
    // Hermitian nb = 2 pack (valid when pd.nrep > 0)
    PairDesc pd{};
    std::vector<double> pcoef;       // PAIR_HEAD + PAIR_MAX*PAIR_REP (when pd.vpaired)
    std::vector<double> scoef;       // PS_HEAD + PAIR_MAX*PS_REP
    std::vector<double> psteps;      // nt * pd.stride
    double lat[6] = {};              // a1, a2 (Angstrom), for the supercell phase of the FFT form
};
// host_state = false leaves rho0 empty: dev_create then computes the initial
// state on the device (same H_k, Jacobi eigensolver and T = 0 occupation).
HostSolve build_solve(const TBModel& m, const SolveParams& p, bool host_state = true);

// The reference's real-space algorithm on the GPU (sbe_fftform.hip): rho(R),
// 2-D FFTs in every RK4 stage, one step replayed as a HIP graph. FP64 only.
struct FFTFormSolve;
FFTFormSolve* fftform_create(const HostSolve& hs);
double fftform_run(FFTFormSolve* fs, double* J_out);  // nt steps; returns device ms
double fftform_last_ms(FFTFormSolve* fs);
void fftform_destroy(FFTFormSolve* fs);
// Unnormalised in-place 2-D FFTs of nch row-major n1 x n2 complex fields on
// the GPU (host buffer in and out), the transform of the FFT form.
void fft2_host(double* data, int nch, int n1, int n2, int sign);
// The same transform on device memory: a plan holds the twiddle tables and a
// stream; exec transforms the nch fields at data (checked to be one device
// allocation large enough) `repeats` times and returns the device time (ms).
struct FFT2Plan;
FFT2Plan* fft2_plan_create(int nch, int n1, int n2);
double fft2_plan_exec(FFT2Plan* p, void* data, int sign, int repeats);
void fft2_plan_destroy(FFT2Plan* p);

// The text dumps of SWESim::test_files (src/swe.cpp:131-181) into test_dir
// and fields_dir; the r <-> k transforms run on the GPU (fft2_host).
void write_test_files(const TBModel& m, const HostSolve& hs, const std::string& test_dir,
                      const std::string& fields_dir);

// Device side (sbe_kernels.hip).
struct DeviceSolve;
DeviceSolve* dev_create(const HostSolve& hs);
// Propagates nt steps from the current device state. J_out: [nt][6], or
// null to leave J(t) in the solve's pinned host buffer only (dev_host_current).
// Returns the device time in milliseconds.
double dev_run(DeviceSolve* ds, int precision, double* J_out);
// J(t) of the last completed run, [nt][6], in the solve's pinned host buffer
// (valid until the next run of the solve).
const double* dev_host_current(const DeviceSolve* ds);
// The same run without allocating device memory: the current partials go to
// the library's static scratch array on the solve's device (at most
// static_scratch_bytes(); larger launches are refused). Concurrent calls of
// this process are serialised.
double dev_run_static(DeviceSolve* ds, int precision, double* J_out);
size_t static_scratch_bytes();
// Asynchronous form: enqueue on the solve's stream, poll, then complete.
void dev_launch(DeviceSolve* ds, int precision);
bool dev_query(DeviceSolve* ds);
double dev_complete(DeviceSolve* ds, double* J_out);
// HBM footprint of a solve built from hs (bytes).
size_t dev_bytes(const HostSolve& hs);
void dev_get_rho(DeviceSolve* ds, double* rho_elem_major);
// Device time of the propagation kernel alone in the last dev_run.
double dev_last_kernel_ms(DeviceSolve* ds);
// 0 = automatic (Hermitian-paired kernel when the blocks allow), 1 = generic.
void dev_set_variant(DeviceSolve* ds, int variant);
// Kernel the next run launches: 1 thread-per-k, 2 Hermitian nb=2, 3 element-per-lane.
int dev_kernel_variant(DeviceSolve* ds, int precision);
// Time the candidate launch configurations of this solve on its first
// `steps` steps (state untouched) and keep the fastest; cached per shape.
// Returns the chosen kernel kind.
int dev_autotune(DeviceSolve* ds, int precision, int steps);
// Process-wide switch: tune every solve's shape on its first run.
void set_autotune(int on);
int get_autotune();
// Launch key the next run uses: kind, a, b, c, fp32.
void dev_launch_key(DeviceSolve* ds, int precision, int* out5);
int dev_kpoints_per_wave(DeviceSolve* ds, int precision);
// Launch key and k-points per lane group under the batch-member rules: out6.
void dev_batch_launch_key(DeviceSolve* ds, int precision, int* out6);
// Launch rules of a solve whose launches share the device with other solves
// (the server's requests): those of batch members.
void dev_set_shared(DeviceSolve* ds, bool shared);
// Flops per k-point and RK4 step of the formulation the next run launches.
double dev_flops_per_kstep(DeviceSolve* ds, int precision);
// Arithmetic per k-point and RK4 step fixed by the model (Hermitian folding,
// live field and current axes), independent of the kernel.
double dev_min_flops_per_kstep(DeviceSolve* ds);
void dev_destroy(DeviceSolve* ds);

// Batched propagation: the given solves advance nt steps in one launch per
// kernel class (different classes overlap on separate streams). A Batch holds
// streams, events and launch tables across runs, not the solves: the caller
// names the members (distinct solves) and their precisions at every run.
struct Batch;
Batch* batch_create();
void batch_destroy(Batch* b);
// J_out[i]: [nt_i][6] for items[i]. Returns device time (ms).
double batch_run(Batch* b, const std::vector<DeviceSolve*>& items, const std::vector<int>& prec,
                 const std::vector<double*>& J_out);

}  // namespace hhg
