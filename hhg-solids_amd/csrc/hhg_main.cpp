// hhg_main.cpp -- standalone driver (the counterpart of src/main.cpp), a
// plain client of the C ABI of libwannier.so (include/hhg_swe.h).
//
//   hhg_swe [path_tb] [key=value ...]
//
// With no arguments it does what src/main.cpp:8-23 does: default Settings
// (include/settings.h), model ../hmcase0_tb.dat, init, one run. Keys: nr1 nr2
// nr3 tmax dt intensity lambda tmax_field (numbers), pol=x,y,z phi=x,y,z,
// precision=fp64|fp32, kernel=auto|generic|fftform, test_files=1 (the dumps of
// src/swe.cpp:131-181 into test/ and fields/), out=FILE (J(t): one line
// "t Re Jx Im Jx Re Jy Im Jy Re Jz Im Jz" per step, t in au).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hhg_swe.h"

static bool parse3(const char* s, double* v) { return std::sscanf(s, "%lf,%lf,%lf", v, v + 1, v + 2) == 3; }

int main(int argc, char** argv) {
    std::string path = "../hmcase0_tb.dat", out, precision = "fp64", kernel = "auto";
    bool dumps = false;
    Settings* st = Settings_new();
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        const char* eq = std::strchr(a, '=');
        if (!eq) { path = a; continue; }
        const std::string key(a, eq - a);
        const char* val = eq + 1;
        double v3[3];
        if (key == "nr1") Settings_set_nr1(st, std::atoi(val));
        else if (key == "nr2") Settings_set_nr2(st, std::atoi(val));
        else if (key == "nr3") Settings_set_nr3(st, std::atoi(val));
        else if (key == "tmax") Settings_set_tmax(st, std::atof(val));
        else if (key == "dt") Settings_set_dt(st, std::atof(val));
        else if (key == "intensity") Settings_set_intensity(st, std::atof(val));
        else if (key == "lambda") Settings_set_lambda(st, std::atof(val));
        else if (key == "tmax_field") Settings_set_tmax_field(st, std::atof(val));
        else if (key == "pol" && parse3(val, v3)) Settings_set_pol_vec(st, v3);
        else if (key == "phi" && parse3(val, v3)) Settings_set_phi_vec(st, v3);
        else if (key == "precision") precision = val;
        else if (key == "kernel") kernel = val;
        else if (key == "test_files") dumps = std::atoi(val) != 0;
        else if (key == "out") out = val;
        else { std::fprintf(stderr, "hhg_swe: unknown argument '%s'\n", a); Settings_delete(st); return 2; }
    }
    SWESim* sim = SWESim_new();
    int rc = 0;
    auto check = [&](const char* what) {
        if (SWESim_get_status(sim) != 0) {
            std::fprintf(stderr, "hhg_swe: %s: %s\n", what, SWESim_get_error(sim));
            rc = 1;
        }
        return rc == 0;
    };
    std::vector<char> pbuf(path.begin(), path.end());
    pbuf.push_back('\0');
    SWESim_set_path_tb(sim, pbuf.data());
    SWESim_set_settings(sim, st);
    SWESim_set_precision(sim, precision == "fp32" ? 1 : precision == "fp64" ? 0 : -1);
    if (check("precision")) {
        SWESim_set_kernel_variant(sim, kernel == "generic" ? 1 : kernel == "fftform" ? 2 : kernel == "auto" ? 0 : -1);
        check("kernel");
    }
    if (rc == 0) {
        SWESim_init(sim);
        check("init");
    }
    if (rc == 0 && dumps) {
        SWESim_test_files(sim);
        check("test_files");
    }
    if (rc == 0) {
        SWESim_run_simulation(sim);
        if (check("run_simulation")) {
            const int nt = SWESim_get_nt(sim);
            std::printf("nt = %d, %d bands, kernel %d, device time %.3f ms\n", nt, SWESim_get_num_orb(sim),
                        SWESim_get_kernel(sim), SWESim_last_run_ms(sim));
            if (!out.empty()) {
                std::vector<double> t(nt);
                std::vector<hhg_cdouble> jx(nt), jy(nt), jz(nt);
                SWESim_get_current(sim, t.data(), jx.data(), jy.data(), jz.data());
                FILE* f = std::fopen(out.c_str(), "w");
                if (!f) { std::fprintf(stderr, "hhg_swe: cannot write %s\n", out.c_str()); rc = 1; }
                else {
                    for (int i = 0; i < nt; ++i)
                        std::fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", t[i], jx[i].real(), jx[i].imag(),
                                     jy[i].real(), jy[i].imag(), jz[i].real(), jz[i].imag());
                    std::fclose(f);
                }
            }
        }
    }
    SWESim_delete(sim);
    Settings_delete(st);
    return rc;
}
