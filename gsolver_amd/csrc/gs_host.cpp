// gs_host.cpp -- host-side set-up of a plan (FP64, once per run).
//
// Follows process_input.cpp:98-182 (load_info), :211-295 (load_nonlindata),
// :297-482 (named parameters), :548-623 (nets, omega grid, k(omega)),
// :677-721 (raw .rawdisp loader), :806-871 (HO dispersion, group velocity),
// :966-1021 (KERR_PROFILE); solver_main.cpp:340-413 (plasma coefficients,
// Raman response); ionization.cpp:313-331 (rate table); fhatha.cpp:177-198
// (QDHT plan); pulse_gg.cpp:6-124 (initial pulse); dumps.cpp:15-45 (dump list).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "gs_host.h"

namespace gs {

namespace {
constexpr double kC = 299792458.0;
constexpr double kMu0 = 1.2566370615e-6;
constexpr double kE = 1.60217733e-19;
constexpr double kMe = 9.1093897e-31;
constexpr double kIonLnStep = 0.003304003148325;
constexpr double kIonMinLn = 29.933606208922594;
constexpr double kIonrateDenomLn = 46.0517018;
constexpr double kAbsWidth = 0.1, kAbsBeta = 10.0;
constexpr int kIonN = 8192;

Error cfg_error(const std::string& m) { return Error(-2, m); }
}  // namespace

// ---------------------------------------------------------------- Config
Config::Config(const std::string& path, const std::vector<std::string>& overrides) : over_(overrides)
{
    std::ifstream f(path);
    if (!f) throw cfg_error("Error opening file with starting condition.");
    std::string t;
    while (f >> t) tok_.push_back(t);
}

bool Config::find(const std::string& name, std::string* val) const
{
    for (size_t i = 0; i + 1 < over_.size(); ++i)
        if (over_[i] == name) { *val = over_[i + 1]; return true; }
    // the file is scanned as "%s %s" pairs; an odd last token pairs with the
    // previous value still held in the reader's buffer
    const std::string* prev = nullptr;
    for (size_t i = 0; i < tok_.size(); i += 2) {
        const std::string* v = (i + 1 < tok_.size()) ? &tok_[i + 1] : prev;
        if (i + 1 < tok_.size()) prev = v;
        if (tok_[i] == name) { *val = v ? *v : std::string(); return true; }
    }
    return false;
}

static std::string notfound(const std::string& n) { return "load_namedparameter: parameter " + n + " not found"; }

double Config::f(const std::string& n) const { std::string v; if (!find(n, &v)) throw cfg_error(notfound(n)); return std::atof(v.c_str()); }
double Config::f(const std::string& n, double d) const { std::string v; return find(n, &v) ? std::atof(v.c_str()) : d; }
int Config::i(const std::string& n) const { std::string v; if (!find(n, &v)) throw cfg_error(notfound(n)); return std::atoi(v.c_str()); }
int Config::i(const std::string& n, int d) const { std::string v; return find(n, &v) ? std::atoi(v.c_str()) : d; }
double Config::numf(const std::string& n, int num, double d) const
{
    std::string v;
    if (find(n + std::to_string(num), &v) || find(n, &v)) return std::atof(v.c_str());
    return d;
}
int Config::numi(const std::string& n, int num, int d) const
{
    std::string v;
    return find(n + std::to_string(num), &v) ? std::atoi(v.c_str()) : d;
}
std::string Config::s(const std::string& n) const { std::string v; if (!find(n, &v)) throw cfg_error(notfound(n)); return v; }
std::string Config::s(const std::string& n, const std::string& d) const { std::string v; return find(n, &v) ? v : d; }
std::string Config::nums(const std::string& n, int num) const { return s(n + std::to_string(num)); }
std::string Config::nums(const std::string& n, int num, const std::string& d) const { return s(n + std::to_string(num), d); }

// ---------------------------------------------------------------- helpers
static std::vector<double> make_net(double a, double b, int n, const std::string& type)
{
    std::vector<double> z(n);
    if (type.empty() || type[0] == 'e') {
        const double step = (b - a) / (n - 1);
        for (int i = 0; i < n; ++i) z[i] = a + i * step;
    } else if (type[0] == 'p') {
        const int pw = std::atoi(type.c_str() + 1);
        auto oddpow = [pw](double x) { return x >= 0 ? std::pow(x, pw) : (pw % 2 == 0 ? NAN : -std::pow(-x, pw)); };
        auto oddroot = [pw](double x) { return x >= 0 ? std::pow(x, 1.0 / pw) : (pw % 2 == 0 ? NAN : -std::pow(-x, 1.0 / pw)); };
        const double n0 = n / (1 - oddroot(b / a));
        const double A = -a / oddpow(n0);
        for (int i = 0; i < n; ++i) z[i] = A * oddpow(i - n0);
    } else {
        throw cfg_error("Unrecognized net type!");
    }
    return z;
}

struct RawDisp {
    std::vector<double> w;
    std::vector<double> n;   // 2 per point (re, im)
};

static int read_disp_type(const std::string& fname, RawDisp* raw)
{
    if (fname.empty() || fname == "NO") return 0;
    std::FILE* f = std::fopen(fname.c_str(), "rb");
    if (!f) throw cfg_error("Unable to open file with refractive index information!");
    int type = 0, np = 0;
    bool ok = std::fread(&type, sizeof(int), 1, f) == 1;
    if (ok && type == 3) {
        ok = std::fread(&np, sizeof(int), 1, f) == 1 && np >= 2;
        if (ok) {
            raw->w.resize(np);
            raw->n.resize(2 * (size_t)np);
            ok = std::fread(raw->w.data(), sizeof(double), np, f) == (size_t)np &&
                 std::fread(raw->n.data(), sizeof(double), 2 * (size_t)np, f) == 2 * (size_t)np;
        }
    }
    std::fclose(f);
    if (!ok) throw cfg_error("Unknown type of file with refractive index information!");
    return type;
}

// load_refrindex_raw: linear interpolation bracketed on signed omega,
// interpolated on |omega|; outside the table n = 1.  Also returns the range.
static void raw_index(const RawDisp& d, double P, const double* w, int n, cd* out, double* wmin, double* wmax)
{
    const int np = (int)d.w.size();
    *wmax = d.w[np - 1];
    *wmin = d.w[0];
    for (int j = 0; j < np; ++j) { *wmax = std::max(*wmax, d.w[j]); *wmin = std::min(*wmin, d.w[j]); }
    for (int i = 0; i < n; ++i) {
        for (int j = 1; j < np; ++j) {
            if ((d.w[j - 1] - w[i]) * (d.w[j] - w[i]) <= 0) {
                const double re0 = P * (d.n[2 * j - 2] - 1) + 1, im0 = P * d.n[2 * j - 1];
                const double re1 = P * (d.n[2 * j] - 1) + 1, im1 = P * d.n[2 * j + 1];
                const double fr = (std::fabs(w[i]) - d.w[j - 1]) / (d.w[j] - d.w[j - 1]);
                out[i] = cd(re0 + (re1 - re0) * fr, im0 + (im1 - im0) * fr);
                break;
            }
            if (j == np - 1) out[i] = 1.0;
        }
    }
}

static void write_file(const std::string& dir, const char* name, const void* p, size_t n)
{
    std::string path = dir + "/" + name;
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return;
    std::fwrite(p, 1, n, f);
    std::fclose(f);
}

// ---------------------------------------------------------------- set-up
Setup build_setup(const Config& cfg, const std::string& workdir)
{
    Setup s;
    // ---- load_info
    const double lambda = cfg.f("LAMBDA_V");
    s.OMEGA0 = 2 * M_PI * kC / lambda;
    const double w0 = s.OMEGA0;
    s.OMEGA_MIN = 0;
    s.OMEGA_MAX = 1e20;
    s.TMIN = cfg.f("T_MIN", NAN);
    s.TMIN = cfg.f("T_MIN_OMEGA", s.TMIN * w0) / w0;
    s.TMAX = cfg.f("T_MAX", NAN);
    s.TMAX = cfg.f("T_MAX_OMEGA", s.TMAX * w0) / w0;
    s.NT = cfg.i("N_T");
    const int nt = s.NT;
    if (nt < 4) throw Error(-1, "N_T must be at least 4");
    s.TSTEP = (s.TMAX - s.TMIN) / nt;
    const std::string dispfile = cfg.s("DISPERSION_FILE");

    // calculate_omega: FFT order around omega0
    s.omega.resize(nt);
    const double wstep = 2 * M_PI / (s.TMAX - s.TMIN);
    for (int i = 0; i < nt; ++i) s.omega[i] = i < nt / 2 ? w0 + wstep * i : w0 - wstep * (nt - i);

    // calculate_wavenumber (for the grid, then for omega0 alone)
    RawDisp raw;
    const int dtype = read_disp_type(dispfile, &raw);
    if (dtype != 0 && dtype != 3)
        throw Error(-1, "dispersion file type not supported by this build (raw type 3 or NO only)");
    const double P = cfg.f("PRESSURE", 1);
    std::vector<cd> ri(nt, cd(1.0, 0.0));
    cd ri0(1.0, 0.0);
    if (dtype == 3) {
        raw_index(raw, P, s.omega.data(), nt, ri.data(), &s.OMEGA_MIN, &s.OMEGA_MAX);
        raw_index(raw, P, &s.OMEGA0, 1, &ri0, &s.OMEGA_MIN, &s.OMEGA_MAX);
    }
    auto wavenumbers = [&](const double* w, const cd* n, int cnt, cd* k) {
        double wmax = w[0];
        for (int i = 0; i < cnt; ++i) wmax = std::max(wmax, w[i]);
        for (int i = 0; i < cnt; ++i) {
            k[i] = std::conj(n[i]) * w[i] / kC;
            const double d = (w[i] / wmax - 1 - kAbsWidth) / kAbsWidth;
            if (d > 0) k[i] += cd(0, -1) * k[i].real() * (std::exp(kAbsBeta) * d * d);
        }
    };
    s.wavenum.resize(nt);
    wavenumbers(s.omega.data(), ri.data(), nt, s.wavenum.data());
    wavenumbers(&s.OMEGA0, &ri0, 1, &s.WAVENUMBER0);
    if (!workdir.empty()) {
        write_file(workdir, "wavenum.bin", s.wavenum.data(), sizeof(cd) * nt);
        write_file(workdir, "omega.bin", s.omega.data(), sizeof(double) * nt);
    }
    // calculate_groupvelocity (kept for the DISPERSION_FILE NO case)
    {
        bool found = false;
        for (int i = 0; i < nt - 1 && !found; ++i)
            if (std::fabs(w0 - s.omega[i]) < wstep) {
                s.GROUP_VELOCITY = (s.omega[i] - s.omega[i + 1]) / (s.wavenum[i] - s.wavenum[i + 1]).real();
                found = true;
            }
        if (!found) throw cfg_error("calculate_groupvelocity: invalid input: omega0 is not inside the omega net");
    }
    // calculate_HOdispersion (raw branch)
    s.ho.assign(nt, 0.0);
    if (cfg.s("DISPERSION_FILE", "NO") != "NO") {
        if (dtype != 3) throw Error(-1, "HO-dispersion calculation from refractive index data of other type than raw is not supported!");
        raw_index(raw, P, s.omega.data(), nt, ri.data(), &s.OMEGA_MIN, &s.OMEGA_MAX);
        double n0, dndw;
        if (w0 != s.omega[0]) {
            n0 = 0;
            dndw = 0;
            for (int i = 0; i < nt - 1; ++i)
                if (std::fabs(w0 - s.omega[i]) < std::fabs(s.omega[1] - s.omega[0])) {
                    dndw = (ri[i + 1] - ri[i]).real() / (s.omega[i + 1] - s.omega[i]);
                    n0 = ri[i].real() + dndw * (w0 - s.omega[i]);
                    break;
                }
        } else {
            n0 = ri[0].real();
            dndw = (ri[1] - ri[nt - 1]).real() / (s.omega[1] - s.omega[nt - 1]);
        }
        const double ngr = n0 + w0 * dndw;
        for (int i = 0; i < nt; ++i) s.ho[i] = s.omega[i] / kC * (ri[i].real() - ngr);
        s.GROUP_VELOCITY = kC / ngr;
        for (int i = 0; i < nt; ++i) {
            const double w = s.omega[i];
            const double dh = (w / s.OMEGA_MAX - 1) / kAbsWidth;
            const double hf = dh < 0 ? 1.0 : std::exp(-kAbsBeta * dh * dh);
            const double dl = (w * (1 - kAbsWidth) / s.OMEGA_MIN - 1);
            const double lf = dl > 0 ? 1.0 : std::exp(-kAbsBeta * dl * dl);
            s.ho[i] *= lf * hf;
        }
    }
    {
        const double clip = (1 - kAbsWidth) * (w0 == s.omega[0] ? s.omega[nt / 2 - 1] : s.omega[nt - 1]);
        s.OMEGA_MAX = std::min(s.OMEGA_MAX, clip);
    }
    s.MAX_TOLERANCE = cfg.f("MAX_TOLERANCE", 1e-4);
    s.NR = cfg.i("N_R");
    if (s.NR < 2) throw Error(-1, "N_R must be at least 2");
    s.XMAX = cfg.f("R_MAX");
    s.NZ = cfg.i("N_Z");
    if (s.NZ < 2) throw Error(-1, "N_Z must be at least 2");
    const double zmin = cfg.f("Z_MIN", 0), zmax = cfg.f("Z_MAX", 0);
    const std::string znt = cfg.s("ZNET_TYPE");
    s.ZNET = make_net(zmin, zmax, s.NZ, znt);
    s.CURRENT_Z0 = s.ZNET[0];
    if (cfg.f("APERTURE_N", 0) != 0) throw Error(-1, "APERTURE_N is not supported by this build");
    const double ambient = cfg.f("AMBIENT_CARRIERS", 0) * cfg.f("NEUTRAL_DENSITY", 0);

    // ---- load_nonlindata
    s.n2 = cfg.f("NONLIN_REFRINDEX", 0) * P;
    if (cfg.f("NONLIN_REFRINDEX4", 0) * P != 0) throw Error(-1, "NONLIN_REFRINDEX4 is not supported by this build");
    s.raman_fraction = cfg.f("RAMAN_FRACTION");
    s.tau_raman = cfg.f("TAU_RAMAN");
    s.omega_raman = cfg.f("OMEGA_RAMAN");
    s.neutral_density = cfg.f("NEUTRAL_DENSITY") * P;
    const double tau_r = cfg.f("RECOMBINATION_TAU");
    s.collision_tau = cfg.f("COLLISION_TAU") / P;
    const double U_joule = cfg.f("IONIZATION_POTENTIAL") * kE;
    const double mred = cfg.f("REDUCED_MASS", 1);
    // ambient carriers: the reference draws rho0 = 2*AMBIENT*drand48() on every
    // call (ionization.cpp:15); this build uses its mean, AMBIENT.
    s.rho_init = ambient;

    // ---- initialize_variables
    const double m = kMe * mred;
    const double k0 = s.WAVENUMBER0.real();
    const double ibs = kE / m * kE * kMu0 / k0;
    const double wt = w0 * s.collision_tau;
    s.avalanche = ibs * wt / (1 + wt * wt);
    s.ponderomotive = ibs / 2.0 / k0 / s.GROUP_VELOCITY / (1.0 + 1.0 / wt);
    s.plasma_func.resize(nt);
    for (int i = 0; i < nt; ++i) {
        const double a = s.collision_tau * s.omega[i];
        const double k = s.wavenum[i].real();
        s.plasma_func[i] = -(a / (1.0 + cd(0, 1) * a) / 2.0 / k * kE * kE / m * kMu0);
    }
    {
        if (cfg.s("IONIZATION_PROFILE", "NO") != "NO") throw Error(-1, "IONIZATION_PROFILE files are not supported by this build");
        s.ionrate_ln.assign(kIonN, 0.0);
        for (int i = 0; i < kIonN - 1; ++i)
            s.ionrate_ln[i] = keldysh_rate_ln(s, kIonMinLn + kIonLnStep * i, U_joule, mred) - kIonrateDenomLn;
    }
    // Raman: R[m] = fR M e^{-m dt/tau} sin(W m dt), RAMAN_FUNCTION = FFT(R) + (1 - fR)
    {
        const double W = s.omega_raman, tau = s.tau_raman, dt = s.TSTEP;
        s.raman_M = dt * (1 + W * W * tau * tau) / tau / (W * tau);
        s.raman_z = std::exp(cd(-dt / tau, W * dt));
        // FFT_fwd(R) + (1 - fR) as a direct DFT on a root table (set-up only)
        std::vector<cd> root(nt);
        for (int k = 0; k < nt; ++k) root[k] = std::exp(cd(0, -2 * M_PI * k / nt));
        std::vector<double> Rm(nt);
        for (int m = 0; m < nt; ++m) Rm[m] = s.raman_fraction * s.raman_M * std::exp(-dt * m / tau) * std::sin(W * dt * m);
        s.raman_function.assign(nt, cd(0, 0));
        for (int k = 0; k < nt; ++k) {
            cd acc(0, 0);
            for (int m = 0; m < nt; ++m) acc += Rm[m] * root[(size_t)k * m % nt];
            s.raman_function[k] = acc + (1 - s.raman_fraction);
        }
    }
    s.ion_potential = U_joule * (1e20 / 1e16);
    s.n2 *= 1e16;
    s.recomb_tau = tau_r * 1e20;
    s.ponderomotive *= 1e20;

    // ---- load_kerr_th_profiles
    if (cfg.s("KERR_PROFILE", "NO") != "NO") throw Error(-1, "KERR_PROFILE files are not supported by this build");
    s.kerr.resize(nt);
    for (int i = 0; i < nt; ++i) s.kerr[i] = s.n2 * s.omega[i] / kC / s.wavenum[i].real() * s.omega[i] / kC;

    // ---- QDHT plan
    {
        const int n = s.NR;
        s.qS = bessel_j0_zero(n + 1);
        s.qV = s.qS / 2.0 / M_PI;
        s.qx.resize(n);
        s.qm1.resize(n);
        for (int i = 0; i < n; ++i) {
            s.qx[i] = bessel_j0_zero(i + 1);
            s.qm1[i] = std::fabs(bessel_j1(s.qx[i]));
        }
        s.qC.resize((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) {
                const double c = 2.0 / s.qS * bessel_j0(s.qx[i] * s.qx[j] / s.qS) / s.qm1[i] / s.qm1[j];
                s.qC[(size_t)i * n + j] = c;
                s.qC[(size_t)j * n + i] = c;
            }
        for (int i = 0; i < n; ++i) s.qx[i] /= s.qS;
    }

    // ---- dump manager (dumps.cpp:15-45)
    {
        const int nd = cfg.i("DUMPS_N", 0);
        const std::string mod = cfg.s("DUMP_FILENAME_MOD", "");
        static const char* known[] = {"full", "maxI", "flux", "plasma_max", "field_axis", "average_spectrum", "plasma_axis"};
        for (int i = 0; i < nd; ++i) {
            DumpSpec d;
            d.type = cfg.nums("DUMPTYPE", i);
            bool ok = false;
            for (const char* k : known) ok |= d.type == k;
            if (!ok) throw Error(-1, "dump type " + d.type + " is not supported by this build");
            const std::string pat = cfg.nums("DUMP_FILENAME", i);
            char buf[1024];
            std::snprintf(buf, sizeof buf, pat.c_str(), mod.c_str());
            d.filename = buf;
            if (d.type == "full") d.nz_denom = cfg.numi("NZ_DENOM", i, 1);
            if (d.type == "maxI" || d.type == "flux") {
                const std::string ft = cfg.nums("FILTER_TYPE", i, "NO");
                d.wfilter.assign(nt, 1.0);
                auto cut = [&](const char* num, bool longpass) {
                    const double wc = 2 * M_PI * kC / std::atof(num);
                    for (int w = 0; w < nt; ++w) d.wfilter[w] = longpass ? (s.omega[w] < wc ? 1.0 : 0.0) : (s.omega[w] > wc ? 1.0 : 0.0);
                };
                if (ft.compare(0, 8, "longpass") == 0) cut(ft.c_str() + 8, true);
                else if (ft.compare(0, 2, "lp") == 0) cut(ft.c_str() + 2, true);
                else if (ft.compare(0, 9, "shortpass") == 0) cut(ft.c_str() + 9, false);
                else if (ft.compare(0, 2, "sp") == 0) cut(ft.c_str() + 2, false);
                else if (ft.compare(0, 2, "NO") != 0) throw Error(-1, "create_wfilter: Unknown filter shape!");
            }
            s.dumps.push_back(d);
        }
    }

    // ---- create_gg (time-domain part; the spectral phase is applied after the FFT)
    {
        if (cfg.s("PULSE_SHAPE").compare(0, 2, "gg") != 0) throw Error(-1, "only PULSE_SHAPE gg is supported by this build");
        double tau_fwhm = cfg.f("DURATION_FWHM", -1);
        tau_fwhm = cfg.f("DURATION_FWHM_OMEGA", tau_fwhm * w0) / w0;
        if (tau_fwhm < 0) throw cfg_error("create_gg(): DURATION_FWHM or DURATION_FWHM_OMGEA should be specified!");
        const double d_fwhm = cfg.f("DIAMETER_FWHM");
        if (cfg.f("NOISE_LEVEL", 0) != 0 || cfg.f("NOISE_LEVEL_UNIFORM", 0) != 0 || cfg.f("NOISE_LEVEL_SPATIAL", 0) != 0)
            throw Error(-1, "unseeded NOISE_LEVEL* keys are not reproducible; use NOISE_MAP_LEVEL and NOISE_SEED");
        const double map_level = cfg.f("NOISE_MAP_LEVEL", 0);
        const uint64_t seed = (uint64_t)cfg.f("NOISE_SEED", 0);
        const double cep = cfg.f("CE_PHASE", 0);
        const double sgT = cfg.f("PULSE_SG_ORDER", 1), sgr = cfg.f("BEAM_SG_ORDER", 1);
        s.phase_flip = cfg.f("PHASE_FLIP", 0) * M_PI;
        s.phase_dip = cfg.f("PHASE_DIP", 0);
        const double pfl = cfg.f("PHASE_FLIP_LAMBDA", 1), pfw = cfg.f("PHASE_FLIP_WIDTH", 0);
        s.pf_maxw = 2 * M_PI * kC / (pfl - pfw / 2.0);
        s.pf_minw = 2 * M_PI * kC / (pfl + pfw / 2.0);
        double E = cfg.f("ENERGY", -1);
        const double epf = 2 * std::pow(std::log(2), 1.0 / 2.0 / sgT) / std::tgamma(1.0 / 2.0 / sgT);
        const double Pcr = 2 * M_PI * kC / k0 / w0 / s.n2 * 1e16;
        double P0 = cfg.f("POWER_CR", E / tau_fwhm * epf / Pcr) * Pcr;
        P0 = cfg.f("POWER", P0);
        E = P0 * tau_fwhm / epf;
        if (E < 0) throw cfg_error("create_gg(): ENERGY or POWER parameter should be specified!");
        const double A0 = std::sqrt(P0 / d_fwhm / d_fwhm * 4 * std::pow(std::log(2), 1.0 / sgr) / M_PI / std::tgamma(1.0 / sgr));
        const double tau = tau_fwhm / 2.0 / std::pow(std::log(std::sqrt(2.0)), 1.0 / 2.0 / sgT);
        const double ro = d_fwhm / 2.0 / std::pow(std::log(std::sqrt(2.0)), 1.0 / 2.0 / sgr);
        const double zR = k0 * ro * ro / 2;
        double C = cfg.f("CHIRP_SI", 0);
        C = cfg.f("CHIRP", C * tau * tau);
        s.chirp_w = cfg.f("CHIRP_SPECTRAL", 0);
        double f = cfg.f("FOCUSING_DISTANCE", 0);
        f = cfg.f("FOCUSING_DISTANCE_DIFFRACTION", f / zR);
        if (zR > 2 * f) f = zR * zR / 2 / f * (1 - std::sqrt(1 - 4 * f * f / zR / zR));
        f = cfg.f("WAVEFRONT_RADIUS", f);
        f = cfg.f("WAVEFRONT_RADIUS_DIFFRACTION", f / zR) * zR;
        const double zmin2 = cfg.f("Z_MIN_DIFFRACTION", s.ZNET[0] / zR) * zR;
        const double zmax2 = cfg.f("Z_MAX_DIFFRACTION", s.ZNET[s.NZ - 1] / zR) * zR;
        s.ZNET = make_net(zmin2, zmax2, s.NZ, znt);

        s.field_t.resize((size_t)s.NR * nt);
        s.init_phase_r.resize(s.NR);
        for (int cx = 0; cx < s.NR; ++cx) {
            const double R = s.qx[cx] * s.XMAX;
            const double amax = A0 * std::exp(-std::pow(R / ro, 2.0 * sgr));
            s.init_phase_r[cx] = R * R / 2.0 / f;
            for (int ct = 0; ct < nt; ++ct) {
                const double t = s.TMIN + s.TSTEP * ct;
                const double T = t / tau;
                cd v = amax * std::exp(-std::pow(T, 2.0 * sgT)) * std::exp(cd(0, cep - T * T * C));
                if (map_level != 0) {
                    // seeded amplitude noise (1 + level*u), u = SplitMix64 -> U[-1,1)
                    uint64_t z = seed + ((uint64_t)cx * nt + ct + 1) * 0x9E3779B97F4A7C15ull;
                    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                    z ^= z >> 31;
                    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
                    v *= 1.0 + map_level * u;
                }
                s.field_t[(size_t)cx * nt + ct] = v;
            }
        }
    }
    return s;
}

cd init_spectral_factor(const Setup& s, int r, int w)
{
    const double k = s.wavenum[w].real();
    const double dw = s.omega[w] - s.OMEGA0;
    double ph = s.init_phase_r[r] * k - s.chirp_w * dw * dw / 2;
    const bool in_band = s.pf_minw < s.omega[w] && s.omega[w] < s.pf_maxw;
    if (s.phase_flip != 0 && in_band) ph += s.phase_flip;
    cd v = std::exp(cd(0, ph));
    if (s.phase_dip != 0 && in_band) v *= (1.0 - s.phase_dip);
    return v;
}

}  // namespace gs
