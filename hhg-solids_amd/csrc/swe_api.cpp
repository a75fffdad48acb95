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
#include <complex>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "sbe_host.h"

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
    Settings* settings = nullptr;
    int precision = 0;
    int variant = 0;
    std::unique_ptr<hhg::HostSolve> host;
    hhg::DeviceSolve* dev = nullptr;
    std::vector<cplx> J[3];  // appended per run, like Observable::_data
    int status = 0;
    std::string err;
    double last_ms = 0.0;
    ~SWESim() { hhg::dev_destroy(dev); }
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

static void sim_init(SWESim* s, bool device = true) {
    if (!s->settings) throw std::runtime_error("SWESim_init: no Settings (call SWESim_set_settings)");
    Settings& st = *s->settings;
    hhg::TBModel m = hhg::read_tb(s->path_tb);
    st.nt = (int)(st.tmax / st.dt);  // src/swe.cpp:52-55
    st.path_tb = s->path_tb;
    st.num_orb = m.nb;
    st.num_sites = m.nR;
    hhg::SolveParams p;
    p.nr1 = st.nr1; p.nr2 = st.nr2; p.nr3 = st.nr3;
    p.tmax = st.tmax; p.dt = st.dt;
    p.lambda_nm = st.lambda_nm; p.tmax_field = st.tmax_field; p.intensity = st.intensity_wcm2;
    p.s[0] = st.sx; p.s[1] = st.sy; p.s[2] = st.sz;
    p.phi[0] = st.phi_x; p.phi[1] = st.phi_y; p.phi[2] = st.phi_z;
    auto hs = std::make_unique<hhg::HostSolve>(hhg::build_solve(m, p));
    hhg::DeviceSolve* dev = device ? hhg::dev_create(*hs) : nullptr;
    if (dev) hhg::dev_set_variant(dev, s->variant);
    hhg::dev_destroy(s->dev);
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
        std::vector<double> J((size_t)6 * nt);
        s->last_ms = hhg::dev_run(s->dev, s->precision, J.data());
        for (int a = 0; a < 3; ++a)
            for (int ti = 0; ti < nt; ++ti) s->J[a].push_back(cplx(J[(size_t)6 * ti + 2 * a], J[(size_t)6 * ti + 2 * a + 1]));
    });
}

// Debug text dumps of src/swe.cpp:131-181 are not part of this library.
API void SWESim_test_files(SWESim* s) {
    guarded(s, [&] { throw std::runtime_error("SWESim_test_files: debug dumps are not provided by this build"); });
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

API void SWESim_delete(SWESim* s) { delete s; }

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
        if (v != 0 && v != 1) throw std::runtime_error("kernel variant must be 0 (automatic) or 1 (generic)");
        s->variant = v;
        if (s->dev) hhg::dev_set_variant(s->dev, v);
    });
}
API double SWESim_flops_per_kstep(SWESim* s) { return (s && s->dev) ? hhg::dev_flops_per_kstep(s->dev) : 0.0; }
API int SWESim_get_kernel(SWESim* s) { return (s && s->dev) ? hhg::dev_kernel_variant(s->dev) : 0; }
API double SWESim_last_kernel_ms(SWESim* s) { return (s && s->dev) ? hhg::dev_last_kernel_ms(s->dev) : 0.0; }

}  // extern "C"
