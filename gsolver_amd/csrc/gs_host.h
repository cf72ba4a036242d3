// gs_host.h -- host-side set-up of a gsolver_solid_qdht run: config lookup,
// grids, dispersion, coefficient tables, Keldysh rate table, QDHT plan and
// the analytic initial field.  Everything here is FP64 and runs once per
// plan; the per-step work runs on the GPU (gs_line.hip, gs_qdht.hip).
#pragma once

#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace gs {

using cd = std::complex<double>;

// Error carrying the reference's message text and a gs_status code.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Named-parameter lookup with the semantics of process_input.cpp:297-482:
// command-line "NAME VALUE" overrides first, then the file read as
// whitespace-separated (name, value) pairs; numbered keys NAME%d.
class Config {
public:
    Config(const std::string& path, const std::vector<std::string>& overrides);
    bool find(const std::string& name, std::string* val) const;
    double f(const std::string& name) const;                    // required
    double f(const std::string& name, double def) const;        // optional
    int i(const std::string& name) const;
    int i(const std::string& name, int def) const;
    double numf(const std::string& name, int num, double def) const;   // NAME%d, then NAME
    int numi(const std::string& name, int num, int def) const;         // NAME%d only
    std::string s(const std::string& name) const;
    std::string s(const std::string& name, const std::string& def) const;
    std::string nums(const std::string& name, int num) const;
    std::string nums(const std::string& name, int num, const std::string& def) const;

private:
    std::vector<std::string> tok_, over_;
};

struct DumpSpec {
    std::string type;       // full, maxI, flux, plasma_max, field_axis, average_spectrum, plasma_axis
    std::string filename;   // after DUMP_FILENAME_MOD substitution
    int nz_denom = 1;
    std::vector<double> wfilter;   // maxI / flux spectral filter (1 or 0 per omega)
};

// Everything the device plan needs, in solver units (after the rescaling of
// solver_main.cpp:385-393).
struct Setup {
    int NR = 0, NT = 0, NZ = 0;
    double TMIN = 0, TMAX = 0, TSTEP = 0, XMAX = 0;
    std::vector<double> ZNET;
    double CURRENT_Z0 = 0;           // CURRENT_Z as set before create_gg rebuilds ZNET
    double OMEGA0 = 0, OMEGA_MIN = 0, OMEGA_MAX = 0, GROUP_VELOCITY = 0;
    cd WAVENUMBER0;
    double MAX_TOLERANCE = 1e-4;
    std::vector<double> omega, ho, kerr;
    std::vector<cd> wavenum, plasma_func, raman_function;

    // nonlinear scalars (solver units)
    double n2 = 0;                   // NONLIN_REFRINDEX * 1e16
    double raman_fraction = 0, tau_raman = 0, omega_raman = 0;
    double raman_M = 0;              // TSTEP (1 + W^2 tau^2) / tau / (W tau)
    cd raman_z;                      // exp((-1/tau + i W) TSTEP)
    double neutral_density = 0, recomb_tau = 0, collision_tau = 0;
    double ion_potential = 0;        // U * 1e4 (joules * IONRATE_DENOM/INTENSITY_DENOM)
    double avalanche = 0, ponderomotive = 0, rho_init = 0;
    std::vector<double> ionrate_ln;  // 8192 entries, [8191] = 0

    // QDHT plan (fhatha.cpp:177-198)
    double qS = 0, qV = 0;
    std::vector<double> qx, qm1, qC;   // qC row-major [i][j]

    // initial field
    std::vector<cd> field_t;         // E(r, t), analytic, FP64 (before the forward FFT)
    std::vector<double> init_phase_r;   // R^2 / (2 f) per radial line (times k(omega))
    double chirp_w = 0;              // CHIRP_SPECTRAL
    double phase_flip = 0, phase_dip = 0, pf_minw = 0, pf_maxw = 0;

    std::vector<DumpSpec> dumps;
};

// Runs load_info, load_nonlindata, the host part of initialize_variables,
// load_kerr_th_profiles, the dump-manager parsing and the time-domain part of
// create_gg.  Writes omega.bin / wavenum.bin into workdir when non-empty.
Setup build_setup(const Config& cfg, const std::string& workdir);

// spectral phase applied by create_gg after the forward FFT (pulse_gg.cpp:107-121)
cd init_spectral_factor(const Setup& s, int r, int w);

// special functions (gs_special.cpp)
double bessel_j0(double x);
double bessel_j1(double x);
double bessel_j0_zero(int k);
double dawson(double x);
void ellipke(double m, double* K, double* E);
double keldysh_rate_ln(const Setup& s, double lnI, double U_joule, double reduced_mass);

}  // namespace gs
