// swe_api.cpp -- the C ABI of libwannier.so (include/hhg_swe.h).
//
// Mirrors src/cwrapper.cpp:5-28 symbol for symbol and the SWESim life cycle
// of src/swe.cpp: init builds the whole solve (src/swe.cpp:49-93; here:
// host tables + device upload), run_simulation propagates nt steps and
// appends nt current samples (src/swe.cpp:183-223, src/observable.cpp:29),
// get_current copies the first nt samples and the time grid in au
// (src/swe.cpp:235-242), restart rebuilds from the file (src/swe.cpp:18-35).
// Unlike the reference, no exception or exit crosses the ABI: errors are
// recorded on the handle and in a thread-local last-error string.
#include <algorithm>
#include <complex>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "hhg_swe.h"
#include "sbe_host.h"
#include "sbe_server.h"

using hhg::cplx;

// Same fields and defaults as include/settings.h:5-31. Opaque to callers.
struct Settings {
    int nr1 = 400, nr2 = 400, nr3 = 1;
    double tmax = 90.0;
    double dt = 10.98e-3 / 0.5;
    int nt = (int)(90.0 / (10.98e-3 / 0.5));
    double lambda_nm = 3000;
    double tmax_field = 80.0;
    double intensity_wcm2 = 5e10;
    double sx = 0.96, sy = 0.25, sz = 0.0;
    double phi_x = 0.0, phi_y = 0.0, phi_z = 0.0;
    int num_orb = 0, num_sites = 0;
    std::string path_tb;
};

struct SWESim {
    std::string path_tb;
    std::vector<hhg_batch*> batches;  // batches this solve is a member of
    Settings* settings = nullptr;
    int precision = 0;
    int variant = 0;
    std::unique_ptr<hhg::HostSolve> host;
    hhg::DeviceSolve* dev = nullptr;
    hhg::FFTFormSolve* ff = nullptr;  // kernel variant 2, created on its first run
    std::vector<cplx> J[3];  // appended per run, like Observable::_data
    int status = 0;
    std::string err;
    double last_ms = 0.0;
    ~SWESim() {
        hhg::dev_destroy(dev);
        hhg::fftform_destroy(ff);
    }
};

// A batch names its members by handle; it resolves their device solves and
// precisions when it runs, so a member's init, restart or set_precision after
// the add is seen by the next batch run, and deleting a member removes it
// from its batches.
struct hhg_batch {
    hhg::Batch* b = nullptr;
    std::vector<SWESim*> sims;
    std::string err;
    double last_ms = 0.0;
    ~hhg_batch() {
        for (SWESim* s : sims) s->batches.erase(std::find(s->batches.begin(), s->batches.end(), this));
        hhg::batch_destroy(b);
    }
};

static thread_local std::string g_last_error;

template <typename F>
static void guarded(SWESim* s, F&& f) {
    try {
        if (!s) throw std::runtime_error("null SWESim handle");
        f();
        s->status = 0;
        s->err.clear();
    } catch (const std::exception& e) {
        g_last_error = e.what();
        if (s) { s->status = 1; s->err = e.what(); }
    } catch (...) {
        g_last_error = "unknown error";
        if (s) { s->status = 1; s->err = g_last_error; }
    }
}

// Appends one run's J(t) ([nt][6] doubles) to the run record. The record
// grows by at least 32 runs at a time (virtual memory until written), so that
// the first runs never copy the earlier ones when it grows.
static void append_run(SWESim* s, const double* J, int nt) {
    for (int a = 0; a < 3; ++a) {
        std::vector<cplx>& v = s->J[a];
        if (v.capacity() < v.size() + nt) v.reserve(std::max(2 * v.capacity(), v.size() + 32 * (size_t)nt));
        for (int ti = 0; ti < nt; ++ti) v.push_back(cplx(J[(size_t)6 * ti + 2 * a], J[(size_t)6 * ti + 2 * a + 1]));
    }
}

static hhg::SolveParams params_of(const Settings& st) {
    hhg::SolveParams p;
    p.nr1 = st.nr1; p.nr2 = st.nr2; p.nr3 = st.nr3;
    p.tmax = st.tmax; p.dt = st.dt;
    p.lambda_nm = st.lambda_nm; p.tmax_field = st.tmax_field; p.intensity = st.intensity_wcm2;
    p.s[0] = st.sx; p.s[1] = st.sy; p.s[2] = st.sz;
    p.phi[0] = st.phi_x; p.phi[1] = st.phi_y; p.phi[2] = st.phi_z;
    return p;
}

static void sim_init(SWESim* s, bool device = true) {
    if (!s->settings) throw std::runtime_error("SWESim_init: no Settings (call SWESim_set_settings)");
    Settings& st = *s->settings;
    hhg::TBModel m = hhg::read_tb(s->path_tb);
    st.nt = (int)(st.tmax / st.dt);  // src/swe.cpp:52-55
    st.path_tb = s->path_tb;
    st.num_orb = m.nb;
    st.num_sites = m.nR;
    const hhg::SolveParams p = params_of(st);
    // device solves compute the initial state on the GPU and keep a host copy
    auto hs = std::make_unique<hhg::HostSolve>(hhg::build_solve(m, p, !device));
    hhg::DeviceSolve* dev = device ? hhg::dev_create(*hs) : nullptr;
    if (dev) {
        hs->rho0.resize((size_t)2 * hs->d.nb * hs->d.nb * hs->d.N);
        hhg::dev_get_rho(dev, hs->rho0.data());
    }
    if (dev) hhg::dev_set_variant(dev, s->variant);
    hhg::dev_destroy(s->dev);
    hhg::fftform_destroy(s->ff);
    s->ff = nullptr;
    s->dev = dev;
    s->host = std::move(hs);
    for (auto& j : s->J) j.clear();
}

extern "C" {

#define API __attribute__((visibility("default")))

API SWESim* SWESim_new() {
    try { return new SWESim(); } catch (...) { g_last_error = "allocation failed"; return nullptr; }
}

API void SWESim_run_simulation(SWESim* s) {
    guarded(s, [&] {
        if (!s->dev) throw std::runtime_error("SWESim_run_simulation: call SWESim_init first (SWESim_init_host sets up the host tables only)");
        const int nt = s->host->d.nt;
        std::vector<double> Jf;
        const double* J = nullptr;  // [nt][6]
        if (s->variant == 2) {  // the reference's real-space FFT algorithm on the GPU
            if (s->precision != 0) throw std::runtime_error("the FFT form runs in FP64 only");
            if (!s->ff) s->ff = hhg::fftform_create(*s->host);
            Jf.resize((size_t)6 * nt);
            s->last_ms = hhg::fftform_run(s->ff, Jf.data());
            J = Jf.data();
        } else {  // read straight from the solve's pinned host buffer
            s->last_ms = hhg::dev_run(s->dev, s->precision, nullptr);
            J = hhg::dev_host_current(s->dev);
        }
        append_run(s, J, nt);
    });
}

// SWESim_run_simulation without device allocation: J(t) of this run goes to
// the caller's J_out ([nt][6] doubles: re/im of x, y, z) instead of the
// handle's current record, the current partials to the library's static
// scratch. Host-only state is refused like run_simulation; the FFT form,
// which allocates its own fields, is refused.
API int SWESim_run_simulation_static(SWESim* s, double* J_out) {
    guarded(s, [&] {
        if (!s->dev) throw std::runtime_error("SWESim_run_simulation_static: call SWESim_init first");
        if (!J_out) throw std::runtime_error("SWESim_run_simulation_static: null output");
        if (s->variant == 2) throw std::runtime_error("SWESim_run_simulation_static: not available for the FFT form");
        s->last_ms = hhg::dev_run_static(s->dev, s->precision, J_out);
    });
    return s ? s->status : -1;
}

// Text dumps of src/swe.cpp:131-181 into test/ and fields/ under the working
// directory (both created when missing); the solve's state is not changed.
API void SWESim_test_files(SWESim* s) {
    guarded(s, [&] {
        if (!s->dev) throw std::runtime_error("SWESim_test_files: call SWESim_init first");
        std::printf("Writing test files\n");
        hhg::write_test_files(hhg::read_tb(s->path_tb), *s->host, "test", "fields");
    });
}

API void SWESim_set_path_tb(SWESim* s, char* path) {
    guarded(s, [&] { s->path_tb = path ? std::string(path) : std::string(); });
}

API void SWESim_set_settings(SWESim* s, Settings* st) {
    guarded(s, [&] { s->settings = st; });
}

API void SWESim_init(SWESim* s) { guarded(s, [&] { sim_init(s); }); }

API void SWESim_restart(SWESim* s) { guarded(s, [&] { sim_init(s); }); }

API void SWESim_init_host(SWESim* s) { guarded(s, [&] { sim_init(s, false); }); }

API void SWESim_get_current(SWESim* s, double* t, cplx* jx, cplx* jy, cplx* jz) {
    guarded(s, [&] {
        if (!s->host) throw std::runtime_error("SWESim_get_current: solve not initialised");
        const int nt = s->host->d.nt;
        cplx* out[3] = {jx, jy, jz};
        for (int ti = 0; ti < nt; ++ti) {
            if (t) t[ti] = s->host->t[ti];
            for (int a = 0; a < 3; ++a)
                if (out[a]) out[a][ti] = ti < (int)s->J[a].size() ? s->J[a][ti] : cplx(0.0, 0.0);
        }
    });
}

API int SWESim_get_current_run(SWESim* s, int run, double* t, cplx* jx, cplx* jy, cplx* jz) {
    guarded(s, [&] {
        if (!s->host) throw std::runtime_error("SWESim_get_current_run: solve not initialised");
        const int nt = s->host->d.nt;
        const int runs = nt > 0 ? (int)(s->J[0].size() / nt) : 0;
        if (run < 0) run += runs;
        if (run < 0 || run >= runs)
            throw std::runtime_error("SWESim_get_current_run: run " + std::to_string(run) + " of " + std::to_string(runs));
        cplx* out[3] = {jx, jy, jz};
        for (int ti = 0; ti < nt; ++ti) {
            if (t) t[ti] = s->host->t[ti];
            for (int a = 0; a < 3; ++a)
                if (out[a]) out[a][ti] = s->J[a][(size_t)run * nt + ti];
        }
    });
    return s ? s->status : -1;
}

API void SWESim_delete(SWESim* s) {
    if (!s) return;
    for (hhg_batch* b : s->batches) b->sims.erase(std::find(b->sims.begin(), b->sims.end(), s));
    delete s;
}

API Settings* Settings_new() {
    try { return new Settings(); } catch (...) { g_last_error = "allocation failed"; return nullptr; }
}
API void Settings_delete(Settings* s) { delete s; }
API void Settings_set_nr1(Settings* s, int v) { if (s) s->nr1 = v; }
API void Settings_set_nr2(Settings* s, int v) { if (s) s->nr2 = v; }
API void Settings_set_nr3(Settings* s, int v) { if (s) s->nr3 = v; }
API void Settings_set_tmax(Settings* s, double v) { if (s) s->tmax = v; }
API void Settings_set_dt(Settings* s, double v) { if (s) s->dt = v; }
API void Settings_set_intensity(Settings* s, double v) { if (s) s->intensity_wcm2 = v; }
API void Settings_set_lambda(Settings* s, double v) { if (s) s->lambda_nm = v; }
API void Settings_set_tmax_field(Settings* s, double v) { if (s) s->tmax_field = v; }
API void Settings_set_pol_vec(Settings* s, double* v) { if (s && v) { s->sx = v[0]; s->sy = v[1]; s->sz = v[2]; } }
API void Settings_set_phi_vec(Settings* s, double* v) { if (s && v) { s->phi_x = v[0]; s->phi_y = v[1]; s->phi_z = v[2]; } }

API int SWESim_get_status(SWESim* s) { return s ? s->status : -1; }
API const char* SWESim_get_error(SWESim* s) { return s ? s->err.c_str() : "null SWESim handle"; }
API const char* hhg_last_error() { return g_last_error.c_str(); }
API int SWESim_get_nt(SWESim* s) { return (s && s->host) ? s->host->d.nt : (s && s->settings ? (int)(s->settings->tmax / s->settings->dt) : 0); }
API int SWESim_get_num_orb(SWESim* s) { return (s && s->host) ? s->host->d.nb : 0; }
API void SWESim_set_precision(SWESim* s, int p) {
    guarded(s, [&] {
        if (p != 0 && p != 1) throw std::runtime_error("precision must be 0 (FP64) or 1 (FP32 stages)");
        s->precision = p;
    });
}
API int SWESim_get_fields(SWESim* s, double* t, double* E, double* A) {
    guarded(s, [&] {
        if (!s->host) throw std::runtime_error("SWESim_get_fields: solve not initialised");
        const int nt = s->host->d.nt;
        for (int i = 0; i < nt; ++i) t[i] = s->host->t[i];
        for (int a = 0; a < 3; ++a)
            for (int i = 0; i < nt; ++i) { E[(size_t)a * nt + i] = s->host->E[a][i]; A[(size_t)a * nt + i] = s->host->A[a][i]; }
    });
    return s ? s->status : -1;
}
API int SWESim_get_rho_k(SWESim* s, cplx* rho) {
    guarded(s, [&] {
        if (!s->host) throw std::runtime_error("SWESim_get_rho_k: solve not initialised");
        const hhg::SolveDesc& d = s->host->d;
        const int nb2 = d.nb * d.nb;
        std::vector<double> em;
        if (s->dev) {  // current device state
            em.resize((size_t)2 * nb2 * d.N);
            hhg::dev_get_rho(s->dev, em.data());
        }
        const std::vector<double>& src = s->dev ? em : s->host->rho0;  // host-only: the initial state
        for (int k = 0; k < d.N; ++k)
            for (int e = 0; e < nb2; ++e)
                rho[(size_t)k * nb2 + e] = cplx(src[2 * ((size_t)e * d.N + k)], src[2 * ((size_t)e * d.N + k) + 1]);
    });
    return s ? s->status : -1;
}
API double SWESim_last_run_ms(SWESim* s) { return s ? s->last_ms : 0.0; }
API void SWESim_set_kernel_variant(SWESim* s, int v) {
    guarded(s, [&] {
        if (v < 0 || v > 2)
            throw std::runtime_error("kernel variant must be 0 (automatic), 1 (generic) or 2 (FFT form)");
        s->variant = v;
        if (s->dev) hhg::dev_set_variant(s->dev, v == 1 ? 1 : 0);
    });
}
API double SWESim_flops_per_kstep(SWESim* s) {
    return (s && s->dev) ? hhg::dev_flops_per_kstep(s->dev, s->precision) : 0.0;
}
API double SWESim_min_flops_per_kstep(SWESim* s) {
    return (s && s->dev) ? hhg::dev_min_flops_per_kstep(s->dev) : 0.0;
}
API int SWESim_get_kernel(SWESim* s) {
    if (s && s->dev && s->variant == 2) return 7;
    return (s && s->dev) ? hhg::dev_kernel_variant(s->dev, s->precision) : 0;
}
API int hhg_fft2(double* data, int nch, int n1, int n2, int sign) {
    try {
        if (!data) throw std::runtime_error("hhg_fft2: null data");
        hhg::fft2_host(data, nch, n1, n2, sign);
        return 0;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return -1;
    }
}
struct hhg_fft2_plan {
    hhg::FFT2Plan* p = nullptr;
    ~hhg_fft2_plan() { hhg::fft2_plan_destroy(p); }
};
API hhg_fft2_plan* hhg_fft2_plan_create(int nch, int n1, int n2) {
    try {
        auto* h = new hhg_fft2_plan();
        try {
            h->p = hhg::fft2_plan_create(nch, n1, n2);
        } catch (...) {
            delete h;
            throw;
        }
        return h;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return nullptr;
    }
}
API double hhg_fft2_plan_exec(hhg_fft2_plan* h, void* data, int sign, int repeats) {
    try {
        if (!h) throw std::runtime_error("hhg_fft2_plan_exec: null plan");
        return hhg::fft2_plan_exec(h->p, data, sign, repeats);
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return -1.0;
    }
}
API void hhg_fft2_plan_destroy(hhg_fft2_plan* h) { delete h; }
API int SWESim_autotune(SWESim* s, int steps) {
    int kind = -1;
    guarded(s, [&] {
        if (!s->dev) throw std::runtime_error("SWESim_autotune: call SWESim_init first");
        if (steps < 1) throw std::runtime_error("SWESim_autotune: steps must be positive");
        kind = hhg::dev_autotune(s->dev, s->precision, steps);
    });
    return s && s->status == 0 ? kind : -1;
}
API int SWESim_get_launch(SWESim* s, int* out5) {
    guarded(s, [&] {
        if (!s->dev) throw std::runtime_error("SWESim_get_launch: call SWESim_init first");
        if (!out5) throw std::runtime_error("SWESim_get_launch: null output");
        hhg::dev_launch_key(s->dev, s->precision, out5);
    });
    return s ? s->status : -1;
}
API int SWESim_get_batch_launch(SWESim* s, int* out6) {
    guarded(s, [&] {
        if (!s->dev) throw std::runtime_error("SWESim_get_batch_launch: call SWESim_init first");
        if (!out6) throw std::runtime_error("SWESim_get_batch_launch: null output");
        hhg::dev_batch_launch_key(s->dev, s->precision, out6);
    });
    return s ? s->status : -1;
}
API int SWESim_get_kpoints_per_wave(SWESim* s) {
    int kpw = -1;
    guarded(s, [&] {
        if (!s->dev) throw std::runtime_error("SWESim_get_kpoints_per_wave: call SWESim_init first");
        kpw = hhg::dev_kpoints_per_wave(s->dev, s->precision);
    });
    return s && s->status == 0 ? kpw : -1;
}
API void hhg_set_autotune(int on) { hhg::set_autotune(on); }
API int hhg_get_autotune() { return hhg::get_autotune(); }
API double SWESim_last_kernel_ms(SWESim* s) { return (s && s->dev) ? hhg::dev_last_kernel_ms(s->dev) : 0.0; }

}  // extern "C"

// ---- batches -----------------------------------------------------------

template <typename F>
static int bguard(hhg_batch* b, F&& f) {
    try {
        if (!b) throw std::runtime_error("null batch handle");
        f();
        b->err.clear();
        return 0;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        if (b) b->err = e.what();
        return 1;
    }
}

extern "C" {

API hhg_batch* hhg_batch_new() {
    try {
        auto* b = new hhg_batch();
        b->b = hhg::batch_create();
        return b;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return nullptr;
    }
}
API void hhg_batch_delete(hhg_batch* b) { delete b; }
API int hhg_batch_add(hhg_batch* b, SWESim* s) {
    return bguard(b, [&] {
        if (!s || !s->dev) throw std::runtime_error("hhg_batch_add: solve not initialised on the device");
        if (s->variant == 2) throw std::runtime_error("hhg_batch_add: the FFT form (kernel variant 2) runs alone");
        if (std::find(b->sims.begin(), b->sims.end(), s) != b->sims.end())
            throw std::runtime_error("hhg_batch_add: solve already in this batch");
        b->sims.push_back(s);
        s->batches.push_back(b);
    });
}
API int hhg_batch_size(hhg_batch* b) { return b ? (int)b->sims.size() : 0; }
API int hhg_batch_run(hhg_batch* b) {
    return bguard(b, [&] {
        std::vector<std::vector<double>> J(b->sims.size());
        std::vector<double*> out(b->sims.size());
        std::vector<hhg::DeviceSolve*> items(b->sims.size());
        std::vector<int> prec(b->sims.size());
        for (size_t i = 0; i < b->sims.size(); ++i) {
            SWESim* s = b->sims[i];
            if (!s->dev || !s->host)
                throw std::runtime_error("hhg_batch_run: member " + std::to_string(i) + " is not initialised on the device");
            if (s->variant == 2)
                throw std::runtime_error("hhg_batch_run: member " + std::to_string(i) + " uses the FFT form (kernel variant 2), which runs alone");
            items[i] = s->dev;
            prec[i] = s->precision;
            J[i].resize((size_t)6 * s->host->d.nt);
            out[i] = J[i].data();
        }
        const double ms = hhg::batch_run(b->b, items, prec, out);
        b->last_ms = ms;
        for (size_t i = 0; i < b->sims.size(); ++i) {
            SWESim* s = b->sims[i];
            append_run(s, J[i].data(), s->host->d.nt);
            s->last_ms = ms;
        }
    });
}
}  // extern "C"

// ---- server ------------------------------------------------------------
struct hhg_server {
    std::unique_ptr<hhg::Server> srv;
};

template <typename F>
static int sguard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return -1;
    }
}

extern "C" {

API hhg_server* hhg_server_new(double hbm_budget_gb, int max_inflight, int host_threads) {
    try {
        hhg::ServerConfig c;
        if (hbm_budget_gb > 0) c.hbm_budget = (size_t)(hbm_budget_gb * (double)(1ull << 30));
        if (max_inflight > 0) c.max_inflight = max_inflight;
        if (host_threads > 0) c.host_threads = host_threads;
        auto* s = new hhg_server();
        s->srv = std::make_unique<hhg::Server>(c);
        return s;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return nullptr;
    }
}
API void hhg_server_delete(hhg_server* s) { delete s; }
API int hhg_server_set_policy(hhg_server* s, int policy, double aging_s) {
    return sguard([&] {
        if (!s) throw std::runtime_error("null server");
        s->srv->set_policy(policy, aging_s);
    });
}
API long long hhg_server_submit(hhg_server* s, const char* path_tb, Settings* st, int precision) {
    long long id = -1;
    sguard([&] {
        if (!s || !st || !path_tb) throw std::runtime_error("hhg_server_submit: null argument");
        if (precision != 0 && precision != 1) throw std::runtime_error("precision must be 0 or 1");
        id = s->srv->submit(path_tb, params_of(*st), precision);
    });
    return id;
}
API int hhg_server_wait(hhg_server* s, long long id, double timeout_s) {
    int st = -1;
    if (sguard([&] { st = s->srv->wait(id, timeout_s); })) return -1;
    return st;
}
API int hhg_server_state(hhg_server* s, long long id) {
    int st = -1;
    if (sguard([&] { st = s->srv->state(id); })) return -1;
    return st;
}
API int hhg_server_nt(hhg_server* s, long long id) {
    int n = -1;
    if (sguard([&] { n = s->srv->nt(id); })) return -1;
    return n;
}
API int hhg_server_result(hhg_server* s, long long id, double* t, cplx* jx, cplx* jy, cplx* jz) {
    return sguard([&] {
        const int nt = s->srv->nt(id);
        std::vector<double> J((size_t)6 * nt);
        s->srv->result(id, t, J.data());
        cplx* out[3] = {jx, jy, jz};
        for (int a = 0; a < 3; ++a)
            if (out[a])
                for (int ti = 0; ti < nt; ++ti) out[a][ti] = cplx(J[(size_t)6 * ti + 2 * a], J[(size_t)6 * ti + 2 * a + 1]);
    });
}
API int hhg_server_timing(hhg_server* s, long long id, double* out7) {
    return sguard([&] { s->srv->timing(id, out7); });
}
API int hhg_server_release(hhg_server* s, long long id) {
    return sguard([&] { s->srv->release(id); });
}
API int hhg_server_stats(hhg_server* s, double* out) {
    return sguard([&] {
        const hhg::ServerStats st = s->srv->stats();
        const double v[10] = {(double)st.submitted, (double)st.completed, (double)st.failed, (double)st.inflight,
                              (double)st.queued, (double)st.peak_inflight, st.ksteps, st.bytes_in_use,
                              st.peak_bytes, 0.0};
        std::copy(v, v + 10, out);
    });
}
API const char* hhg_server_error(hhg_server* s, long long id) {
    static thread_local std::string msg;
    msg.clear();
    if (sguard([&] { msg = s->srv->error(id); })) msg = g_last_error;
    return msg.c_str();
}

API double hhg_batch_last_ms(hhg_batch* b) { return b ? b->last_ms : 0.0; }
API const char* hhg_batch_get_error(hhg_batch* b) { return b ? b->err.c_str() : "null batch handle"; }

}  // extern "C"
