/* o_setup.c -- configuration, grids, dispersion, coefficient tables, QDHT plan
 * and initial field for the oracle (test infrastructure only).
 *
 * Restates process_input.cpp:33-182 (process_input, load_info),
 * :211-295 (load_nonlindata), :297-482 (named-parameter lookup),
 * :548-582 (create_net, calculate_omega), :585-623,677-721 (k(omega) and the
 * raw .rawdisp loader), :734-753 (init_zstep_kerr), :806-871 (HO dispersion),
 * :966-1021 (KERR_PROFILE); solver_main.cpp:265-413 (initialize_variables,
 * calc_raman_function); ionization.cpp:203-232,313-331 (Keldysh table);
 * fhatha.cpp:177-198 (QDHT plan); pulse_gg.cpp:6-124 (create_gg).
 */
#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "o_internal.h"

/* -------------------------------------------------------------------------- */
/* named parameters: argv overrides first, then the file read as "%s %s" pairs */

static const char* over_lookup(orc_sim* s, const char* name)
{
    for (int i = 0; i + 1 < s->cfg.nover; ++i)
        if (strcmp(s->cfg.over[i], name) == 0) return s->cfg.over[i + 1];
    return NULL;
}

static const char* file_lookup(orc_sim* s, const char* name)
{
    /* process_input.cpp:305-312: fscanf("%s %s") in pairs until EOF; a final
     * odd token is compared with the previous pair's value still in valbuf */
    const char* val = NULL;
    for (int i = 0; i < s->cfg.ntok; i += 2) {
        const char* nm = s->cfg.tok[i];
        const char* v = (i + 1 < s->cfg.ntok) ? s->cfg.tok[i + 1] : val;
        if (i + 1 < s->cfg.ntok) val = v;
        if (strcmp(nm, name) == 0) return v ? v : "";
    }
    return NULL;
}

int o_cfg_find(orc_sim* s, const char* name, const char** val)
{
    const char* v = over_lookup(s, name);
    if (!v) v = file_lookup(s, name);
    if (v) { *val = v; return 1; }
    return 0;
}

static void missing(orc_sim* s, const char* name)
{
    snprintf(s->err, sizeof s->err, "load_namedparameter: parameter %.400s not found", name);
    o_throw(s, s->err);
}

double o_load_float(orc_sim* s, const char* name, int has_def, double def)
{
    const char* v;
    if (o_cfg_find(s, name, &v)) return atof(v);
    if (!has_def) missing(s, name);
    return def;
}

int o_load_int(orc_sim* s, const char* name, int has_def, int def)
{
    const char* v;
    if (o_cfg_find(s, name, &v)) return atoi(v);
    if (!has_def) missing(s, name);
    return def;
}

double o_load_numfloat(orc_sim* s, const char* name, int num, int has_def, double def)
{
    char nn[600];
    snprintf(nn, sizeof nn, "%s%d", name, num);
    const char* v;
    if (o_cfg_find(s, nn, &v)) return atof(v);
    if (o_cfg_find(s, name, &v)) return atof(v);
    if (!has_def) missing(s, nn);
    return def;
}

int o_load_numint(orc_sim* s, const char* name, int num, int has_def, int def)
{
    char nn[600];
    snprintf(nn, sizeof nn, "%s%d", name, num);
    const char* v;
    if (o_cfg_find(s, nn, &v)) return atoi(v);
    if (!has_def) missing(s, nn);
    return def;
}

void o_load_string(orc_sim* s, const char* name, char* out, int n, int has_def, const char* def)
{
    const char* v;
    if (o_cfg_find(s, name, &v)) { strncpy(out, v, n); out[n - 1] = 0; return; }
    if (!has_def) missing(s, name);
    strncpy(out, def, n);
    out[n - 1] = 0;
}

void o_load_numstring(orc_sim* s, const char* name, int num, char* out, int n, int has_def, const char* def)
{
    char nn[600];
    snprintf(nn, sizeof nn, "%s%d", name, num);
    o_load_string(s, nn, out, n, has_def, def);
}

/* -------------------------------------------------------------------------- */

static void create_net(orc_sim* s, double xmin, double xmax, int n, const char* type, double* net)
{
    if (type == NULL || type[0] == 'e') {
        double step = (xmax - xmin) / (n - 1);
        for (int i = 0; i < n; ++i) net[i] = xmin + i * step;
    } else if (type[0] == 'p') {
        int pw = atoi(type + 1);
        double r = xmax / xmin;
        double root = r >= 0 ? pow(r, 1.0 / pw) : ((pw % 2 == 0) ? NAN : -pow(-r, 1.0 / pw));
        double n0 = n / (1 - root);
        double op = n0 >= 0 ? pow(n0, pw) : ((pw % 2 == 0) ? NAN : -pow(-n0, pw));
        double A = -xmin / op;
        for (int i = 0; i < n; ++i) {
            double x = i - n0;
            double v = x >= 0 ? pow(x, pw) : ((pw % 2 == 0) ? NAN : -pow(-x, pw));
            net[i] = A * v;
        }
    } else {
        o_throw(s, "Unrecognized net type!");
    }
}

static void calculate_omega(int nt, double tmin, double tmax, double w0, double* om)
{
    double wstep = 2 * M_PI / (tmax - tmin);
    for (int i = 0; i < nt / 2; ++i) om[i] = w0 + wstep * i;
    for (int i = nt / 2; i < nt; ++i) om[i] = w0 - wstep * (nt - i);
}

/* process_input.cpp:677-721 */
static void load_refrindex_raw(orc_sim* s, FILE* fid, int n, const double* cf, double _Complex* ri)
{
    double P = o_load_float(s, "PRESSURE", 1, 1);
    int np = 0;
    if (fread(&np, sizeof(int), 1, fid) != 1 || np < 2) o_throw(s, "bad raw dispersion file");
    double* pw = (double*)malloc(sizeof(double) * np);
    double* pn = (double*)malloc(sizeof(double) * 2 * np);
    if (fread(pw, sizeof(double), np, fid) != (size_t)np || fread(pn, sizeof(double), 2 * (size_t)np, fid) != 2 * (size_t)np) {
        free(pw); free(pn);
        o_throw(s, "truncated raw dispersion file");
    }
    s->OMEGA_MAX = pw[np - 1];
    s->OMEGA_MIN = pw[0];
    for (int j = 0; j < np; ++j) {
        if (pw[j] > s->OMEGA_MAX) s->OMEGA_MAX = pw[j];
        if (pw[j] < s->OMEGA_MIN) s->OMEGA_MIN = pw[j];
    }
    for (int i = 0; i < n; ++i) {
        for (int j = 1; j < np; ++j) {
            if ((pw[j - 1] - cf[i]) * (pw[j] - cf[i]) <= 0) {
                double ren_ = P * (pn[2 * j - 2] - 1) + 1, imn_ = P * pn[2 * j - 1];
                double ren = P * (pn[2 * j] - 1) + 1, imn = P * pn[2 * j + 1];
                double renc = ren_ + (ren - ren_) * (fabs(cf[i]) - pw[j - 1]) / (pw[j] - pw[j - 1]);
                double imnc = imn_ + (imn - imn_) * (fabs(cf[i]) - pw[j - 1]) / (pw[j] - pw[j - 1]);
                ri[i] = renc + imnc * I;
                break;
            }
            if (j == np - 1) ri[i] = 1.0;
        }
    }
    free(pw);
    free(pn);
}

static int open_disp(orc_sim* s, const char* fname, FILE** fid)
{
    *fid = NULL;
    if (fname[0] == 0 || strcmp(fname, "NO") == 0) return 0;
    *fid = fopen(fname, "rb");
    if (!*fid) o_throw(s, "Unable to open file with refractive index information!");
    int type = 0;
    if (fread(&type, sizeof(int), 1, *fid) != 1) o_throw(s, "bad dispersion file");
    return type;
}

/* process_input.cpp:585-623 */
static void calculate_wavenumber(orc_sim* s, const char* fname, int n, const double* om, double _Complex* k)
{
    FILE* fid;
    int type = open_disp(s, fname, &fid);
    double _Complex* ri = (double _Complex*)malloc(sizeof(double _Complex) * n);
    if (type == 0) {
        for (int i = 0; i < n; ++i) ri[i] = 1.0;
    } else if (type == 3) {
        load_refrindex_raw(s, fid, n, om, ri);
    } else {
        if (fid) fclose(fid);
        free(ri);
        o_throw(s, "dispersion file type not supported by this build (raw type 3 or NO only)");
    }
    double wmax = om[0];
    for (int i = 0; i < n; ++i) if (om[i] > wmax) wmax = om[i];
    for (int i = 0; i < n; ++i) {
        k[i] = conj(ri[i]) * om[i] / O_LIGHT_VELOCITY;
        double d = (om[i] / wmax - 1 - O_ABS_LAYER_WIDTH) / O_ABS_LAYER_WIDTH;
        if (d > 0) k[i] += -I * creal(k[i]) * (exp(O_ABS_LAYER_BETA) * d * d);
    }
    free(ri);
    if (fid) fclose(fid);
}

static double hfgauss(orc_sim* s, double w)
{
    double d = (w / s->OMEGA_MAX - 1) / O_ABS_LAYER_WIDTH;
    return d < 0 ? 1.0 : exp(-O_ABS_LAYER_BETA * d * d);
}

static double lfgauss(orc_sim* s, double w)
{
    double d = (w * (1 - O_ABS_LAYER_WIDTH) / s->OMEGA_MIN - 1);
    return d > 0 ? 1.0 : exp(-O_ABS_LAYER_BETA * d * d);
}

/* process_input.cpp:806-871 (raw branch) */
static void calculate_ho(orc_sim* s)
{
    char fname[300];
    o_load_string(s, "DISPERSION_FILE", fname, 300, 1, "NO");
    int nt = s->NT;
    if (strncmp(fname, "NO", 300) == 0) {
        for (int i = 0; i < nt; ++i) s->HO_d[i] = 0.0;
        return;
    }
    FILE* fid;
    int type = open_disp(s, fname, &fid);
    if (type != 3) {
        if (fid) fclose(fid);
        o_throw(s, "HO-dispersion calculation from refractive index data of other type than raw is not supported!");
    }
    double _Complex* ri = (double _Complex*)malloc(sizeof(double _Complex) * nt);
    load_refrindex_raw(s, fid, nt, s->OMEGA_d, ri);
    fclose(fid);
    double* om = s->OMEGA_d;
    double wstep = fabs(om[1] - om[0]);
    double n0 = 0, dndw = 0;
    if (s->OMEGA0 != om[0]) {
        for (int i = 0; i < nt - 1; ++i)
            if (fabs(s->OMEGA0 - om[i]) < wstep) {
                dndw = creal(ri[i + 1] - ri[i]) / (om[i + 1] - om[i]);
                n0 = creal(ri[i]) + dndw * (s->OMEGA0 - om[i]);
                break;
            }
    } else {
        n0 = creal(ri[0]);
        dndw = creal(ri[1] - ri[nt - 1]) / (om[1] - om[nt - 1]);
    }
    double ngr = n0 + s->OMEGA0 * dndw;
    for (int i = 0; i < nt; ++i) s->HO_d[i] = om[i] / O_LIGHT_VELOCITY * (creal(ri[i]) - ngr);
    s->GROUP_VELOCITY = O_LIGHT_VELOCITY / ngr;
    free(ri);
    for (int i = 0; i < nt; ++i) {
        double f = lfgauss(s, om[i]) * hfgauss(s, om[i]);
        s->HO_d[i] *= f;
    }
}

/* ionization.cpp:203-232 */
double o_keldysh_rate_ln(orc_sim* s, double lnI)
{
    double k0 = creal(s->WAVENUMBER0);
    double U = s->IONIZATION_POTENTIAL;   /* joules at table-build time */
    double w0 = s->OMEGA0;
    double phase_v = w0 / k0;
    double E = exp(lnI / 2) * sqrt(2 * phase_v / s->GROUP_VELOCITY * phase_v * O_VACUUM_PERMEABILITY);
    double m = O_ELECTRON_MASS * s->REDUCED_MASS;
    double kg = w0 * sqrt(m / O_ELECTRON_CHARGE / O_ELECTRON_CHARGE * U) / E;
    double G = kg * kg / (1 + kg * kg);
    double X = 1 / (1 + kg * kg);
    double Kg, Eg, Kx, Ex;
    o_ellipke(G, &Kg, &Eg);
    o_ellipke(X, &Kx, &Ex);
    double alpha = M_PI * (Kg - Eg) / Ex, beta = M_PI * M_PI / 4.0 / Kx / Ex;
    double x = 2 / M_PI * U / O_HBAR / w0 * Ex / sqrt(G);
    double nu = floor(x + 1) - x;
    int sumN = (int)ceil(10 / alpha);
    if (sumN > 1000) sumN = 1000;
    double Q = 0;
    for (int i = 0; i < sumN; ++i) Q += exp(-i * alpha) * o_dawson(sqrt(beta * (i + 2 * nu)));
    Q = Q * sqrt(M_PI / 2 / Kx);
    return log(2 * w0 / 9 / M_PI) + 1.5 * log(w0 * m / O_HBAR) - 0.75 * log(G) + log(Q) - alpha * floor(x + 1);
}

/* fhatha.cpp:177-198 */
static void qdht_plan(orc_sim* s)
{
    int n = s->NR;
    s->qS = o_j0_zero(n + 1);
    s->qV = s->qS / 2.0 / M_PI;
    s->qx = (double*)malloc(sizeof(double) * n);
    s->qm1 = (double*)malloc(sizeof(double) * n);
    s->qC = (double*)malloc(sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; ++i) {
        s->qx[i] = o_j0_zero(i + 1);
        s->qm1[i] = fabs(o_j1(s->qx[i]));
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            s->qC[i + (size_t)n * j] = 2.0 / s->qS * o_j0(s->qx[i] * s->qx[j] / s->qS) / s->qm1[i] / s->qm1[j];
    for (int i = 0; i < n; ++i) s->qx[i] /= s->qS;
    s->qC_r = (real_t*)malloc(sizeof(real_t) * (size_t)n * n);
    s->qm1_r = (real_t*)malloc(sizeof(real_t) * n);
    for (size_t i = 0; i < (size_t)n * n; ++i) s->qC_r[i] = (real_t)s->qC[i];
    for (int i = 0; i < n; ++i) s->qm1_r[i] = (real_t)s->qm1[i];
}

/* Deterministic amplitude noise map (1 + level*u), u ~ U[-1,1): the seeded
 * replacement for the drand48 noise of pulse_gg.cpp:92-98 (which is seeded by
 * wall-clock time, solver_main.cpp:188).  u is SplitMix64 (Steele, Lea, Flood
 * 2014) of seed + index*golden-gamma, index = r*N_T + t. */
static double noise_u(uint64_t seed, uint64_t idx)
{
    uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

/* pulse_gg.cpp:6-124 for sg orders, chirp, wavefront; PHASE_FLIP/DIP kept */
static void create_gg(orc_sim* s)
{
    double w0 = s->OMEGA0;
    double tau_fwhm = o_load_float(s, "DURATION_FWHM", 1, -1);
    tau_fwhm = o_load_float(s, "DURATION_FWHM_OMEGA", 1, tau_fwhm * w0) / w0;
    if (tau_fwhm < 0) o_throw(s, "create_gg(): DURATION_FWHM or DURATION_FWHM_OMGEA should be specified!");
    double d_fwhm = o_load_float(s, "DIAMETER_FWHM", 0, 0);
    double nl = o_load_float(s, "NOISE_LEVEL", 1, 0);
    double nlu = o_load_float(s, "NOISE_LEVEL_UNIFORM", 1, 0);
    double nls = o_load_float(s, "NOISE_LEVEL_SPATIAL", 1, 0);
    if (nl != 0 || nlu != 0 || nls != 0)
        o_throw(s, "unseeded NOISE_LEVEL* keys are not reproducible; use NOISE_MAP_LEVEL and NOISE_SEED");
    double map_level = o_load_float(s, "NOISE_MAP_LEVEL", 1, 0);
    uint64_t seed = (uint64_t)o_load_float(s, "NOISE_SEED", 1, 0);
    double cep = o_load_float(s, "CE_PHASE", 1, 0);
    double sgT = o_load_float(s, "PULSE_SG_ORDER", 1, 1);
    double sgr = o_load_float(s, "BEAM_SG_ORDER", 1, 1);
    double phase_flip = o_load_float(s, "PHASE_FLIP", 1, 0) * M_PI;
    double phase_dip = o_load_float(s, "PHASE_DIP", 1, 0);
    double pfl = o_load_float(s, "PHASE_FLIP_LAMBDA", 1, 1);
    double pfw = o_load_float(s, "PHASE_FLIP_WIDTH", 1, 0);
    double pf_maxw = 2 * M_PI * O_LIGHT_VELOCITY / (pfl - pfw / 2.0);
    double pf_minw = 2 * M_PI * O_LIGHT_VELOCITY / (pfl + pfw / 2.0);

    double E = o_load_float(s, "ENERGY", 1, -1);
    double epf = 2 * pow(log(2), 1.0 / 2.0 / sgT) / tgamma(1.0 / 2.0 / sgT);
    double k0 = creal(s->WAVENUMBER0);
    double Pcr = 2 * M_PI * O_LIGHT_VELOCITY / k0 / w0 / s->NONLIN_REFRINDEX * O_INTENSITY_DENOM;
    double P0 = o_load_float(s, "POWER_CR", 1, E / tau_fwhm * epf / Pcr) * Pcr;
    P0 = o_load_float(s, "POWER", 1, P0);
    E = P0 * tau_fwhm / epf;
    if (E < 0) o_throw(s, "create_gg(): ENERGY or POWER parameter should be specified!");
    double A0 = sqrt(P0 / d_fwhm / d_fwhm * 4 * pow(log(2), 1.0 / sgr) / M_PI / tgamma(1.0 / sgr));
    double tau = tau_fwhm / 2.0 / pow(log(sqrt(2.0)), 1.0 / 2.0 / sgT);
    double ro = d_fwhm / 2.0 / pow(log(sqrt(2.0)), 1.0 / 2.0 / sgr);
    double zR = k0 * ro * ro / 2;
    double C = o_load_float(s, "CHIRP_SI", 1, 0);
    C = o_load_float(s, "CHIRP", 1, C * tau * tau);
    double Cw = o_load_float(s, "CHIRP_SPECTRAL", 1, 0);
    double f = o_load_float(s, "FOCUSING_DISTANCE", 1, 0);
    f = o_load_float(s, "FOCUSING_DISTANCE_DIFFRACTION", 1, f / zR);
    if (zR > 2 * f) f = zR * zR / 2 / f * (1 - sqrt(1 - 4 * f * f / zR / zR));
    f = o_load_float(s, "WAVEFRONT_RADIUS", 1, f);
    f = o_load_float(s, "WAVEFRONT_RADIUS_DIFFRACTION", 1, f / zR) * zR;
    double zmin = o_load_float(s, "Z_MIN_DIFFRACTION", 1, s->ZNET[0] / zR) * zR;
    double zmax = o_load_float(s, "Z_MAX_DIFFRACTION", 1, s->ZNET[s->NZ - 1] / zR) * zR;
    char znt[50];
    o_load_string(s, "ZNET_TYPE", znt, 50, 0, "");
    create_net(s, zmin, zmax, s->NZ, znt, s->ZNET);

    int nt = s->NT;
    cplx_t* et = s->lb[0].field_real;
    cplx_t* sp = s->lb[0].b1;
    for (int cx = 0; cx < s->NR; ++cx) {
        double R = s->qx[cx] * s->XMAX;
        double amax = A0 * exp(-pow(R / ro, 2.0 * sgr));
        for (int ct = 0; ct < nt; ++ct) {
            double t = s->TMIN + s->TSTEP * ct;
            double T = t / tau;
            double cp = cep - T * T * C;   /* + 0: OMEGA0 == OMEGA[0] on this grid */
            double _Complex v = amax * exp(-pow(T, 2.0 * sgT)) * cexp(I * cp);
            if (map_level != 0) v *= 1.0 + map_level * noise_u(seed, (uint64_t)cx * nt + ct);
            et[ct] = (cplx_t)v;
        }
        o_fft_exec(&s->fft, et, sp, -1);
        for (int cw = 0; cw < nt; ++cw) {
            double k = creal(s->WAVENUMBER_d[cw]);
            double dw = s->OMEGA_d[cw] - w0;
            double ph = (R * R * k / 2.0 / f) - Cw * dw * dw / 2;
            if (phase_flip != 0 && pf_minw < s->OMEGA_d[cw] && s->OMEGA_d[cw] < pf_maxw) ph += phase_flip;
            double _Complex v = (double _Complex)sp[cw] * cexp(I * ph);
            if (phase_dip != 0 && pf_minw < s->OMEGA_d[cw] && s->OMEGA_d[cw] < pf_maxw) v *= (1.0 - phase_dip);
            s->FIELD[(size_t)cx * nt + cw] = (cplx_t)v;
        }
    }
}

static void write_bin(const char* dir, const char* name, const void* p, size_t sz)
{
    char path[1024];
    snprintf(path, sizeof path, "%s/%s", dir, name);
    FILE* f = fopen(path, "wb");
    if (!f) return;
    fwrite(p, 1, sz, f);
    fclose(f);
}

void o_setup(orc_sim* s, const char* workdir)
{
    /* ---- load_info ---- */
    s->LAMBDA_V = o_load_float(s, "LAMBDA_V", 0, 0);
    s->OMEGA0 = 2 * M_PI * O_LIGHT_VELOCITY / s->LAMBDA_V;
    s->OMEGA_MIN = 0;
    s->OMEGA_MAX = 1e20;
    double w0 = s->OMEGA0;
    s->TMIN = o_load_float(s, "T_MIN", 1, NAN);
    s->TMIN = o_load_float(s, "T_MIN_OMEGA", 1, s->TMIN * w0) / w0;
    s->TMAX = o_load_float(s, "T_MAX", 1, NAN);
    s->TMAX = o_load_float(s, "T_MAX_OMEGA", 1, s->TMAX * w0) / w0;
    s->NT = o_load_int(s, "N_T", 0, 0);
    if (s->NT < 4) o_throw(s, "N_T too small");
    int nt = s->NT;
    s->TSTEP = (s->TMAX - s->TMIN) / nt;
    s->OMEGA_d = (double*)malloc(sizeof(double) * nt);
    s->WAVENUMBER_d = (double _Complex*)malloc(sizeof(double _Complex) * nt);
    s->HO_d = (double*)malloc(sizeof(double) * nt);
    char disp[300];
    o_load_string(s, "DISPERSION_FILE", disp, 300, 0, "");
    calculate_omega(nt, s->TMIN, s->TMAX, w0, s->OMEGA_d);
    calculate_wavenumber(s, disp, nt, s->OMEGA_d, s->WAVENUMBER_d);
    calculate_wavenumber(s, disp, 1, &s->OMEGA0, &s->WAVENUMBER0);
    if (workdir) {
        write_bin(workdir, "wavenum.bin", s->WAVENUMBER_d, sizeof(double _Complex) * nt);
        write_bin(workdir, "omega.bin", s->OMEGA_d, sizeof(double) * nt);
    }
    /* calculate_groupvelocity, process_input.cpp:724-731 */
    {
        double wstep = fabs(s->OMEGA_d[1] - s->OMEGA_d[0]);
        int found = 0;
        for (int i = 0; i < nt - 1; ++i)
            if (fabs(w0 - s->OMEGA_d[i]) < wstep) {
                s->GROUP_VELOCITY = (s->OMEGA_d[i] - s->OMEGA_d[i + 1]) / creal(s->WAVENUMBER_d[i] - s->WAVENUMBER_d[i + 1]);
                found = 1;
                break;
            }
        if (!found) o_throw(s, "calculate_groupvelocity: invalid input: omega0 is not inside the omega net");
    }
    calculate_ho(s);
    if (w0 == s->OMEGA_d[0]) {
        double c = (1 - O_ABS_LAYER_WIDTH) * s->OMEGA_d[nt / 2 - 1];
        if (c < s->OMEGA_MAX) s->OMEGA_MAX = c;
    } else {
        double c = (1 - O_ABS_LAYER_WIDTH) * s->OMEGA_d[nt - 1];
        if (c < s->OMEGA_MAX) s->OMEGA_MAX = c;
    }
    s->MAX_TOLERANCE = o_load_float(s, "MAX_TOLERANCE", 1, O_MAX_TOL_DEFAULT);
    s->NR = o_load_int(s, "N_R", 0, 0);
    if (s->NR < 2) o_throw(s, "N_R too small");
    s->XMAX = o_load_float(s, "R_MAX", 0, 0);
    s->NZ = o_load_int(s, "N_Z", 0, 0);
    if (s->NZ < 2) o_throw(s, "N_Z must be at least 2");
    double zmin = o_load_float(s, "Z_MIN", 1, 0), zmax = o_load_float(s, "Z_MAX", 1, 0);
    char znt[50];
    o_load_string(s, "ZNET_TYPE", znt, 50, 0, "");
    s->ZNET = (double*)malloc(sizeof(double) * s->NZ);
    create_net(s, zmin, zmax, s->NZ, znt, s->ZNET);
    s->CURRENT_Z = s->ZNET[0];
    if (o_load_float(s, "APERTURE_N", 1, 0) != 0) o_throw(s, "APERTURE_N is not supported by this build");
    s->AMBIENT_CARRIER_DENSITY = o_load_float(s, "AMBIENT_CARRIERS", 1, 0) * o_load_float(s, "NEUTRAL_DENSITY", 1, 0);

    /* ---- load_nonlindata ---- */
    double P = o_load_float(s, "PRESSURE", 1, 1);
    s->NONLIN_REFRINDEX = o_load_float(s, "NONLIN_REFRINDEX", 1, 0) * P;
    if (o_load_float(s, "NONLIN_REFRINDEX4", 1, 0) * P != 0) o_throw(s, "NONLIN_REFRINDEX4 is not supported by this build");
    s->RAMAN_FRACTION = o_load_float(s, "RAMAN_FRACTION", 0, 0);
    s->TAU_RAMAN = o_load_float(s, "TAU_RAMAN", 0, 0);
    s->OMEGA_RAMAN = o_load_float(s, "OMEGA_RAMAN", 0, 0);
    s->NEUTRAL_DENSITY = o_load_float(s, "NEUTRAL_DENSITY", 0, 0) * P;
    s->RECOMBINATION_TAU = o_load_float(s, "RECOMBINATION_TAU", 0, 0);
    s->COLLISION_TAU = o_load_float(s, "COLLISION_TAU", 0, 0) / P;
    s->IONIZATION_POTENTIAL = o_load_float(s, "IONIZATION_POTENTIAL", 0, 0) * O_ELECTRON_CHARGE;
    s->REDUCED_MASS = o_load_float(s, "REDUCED_MASS", 1, 1);

    /* ---- initialize_variables ---- */
    int nr = s->NR;
    o_fft_plan_init(&s->fft, nt);
    s->lb = (o_linebuf*)calloc(s->nthreads, sizeof(o_linebuf));
    for (int t = 0; t < s->nthreads; ++t) o_linebuf_alloc(&s->lb[t], nt);
    qdht_plan(s);
    size_t n = (size_t)nr * nt;
    s->FIELD = (cplx_t*)calloc(n, sizeof(cplx_t));
    s->BIG1 = (cplx_t*)calloc(n, sizeof(cplx_t));
    s->BIG2 = (cplx_t*)calloc(n, sizeof(cplx_t));

    double m = O_ELECTRON_MASS * s->REDUCED_MASS;
    double k0 = creal(s->WAVENUMBER0);
    double ibs = O_ELECTRON_CHARGE / m * O_ELECTRON_CHARGE * O_VACUUM_PERMEABILITY / k0;
    double wt = w0 * s->COLLISION_TAU;
    s->AVALANCHE_CROSSSECTION = ibs * wt / (1 + wt * wt);
    s->PONDEROMOTIVE_COEFFICIENT = ibs / 2.0 / k0 / s->GROUP_VELOCITY / (1.0 + 1.0 / wt);
    s->PLASMA_FUNC = (cplx_t*)malloc(sizeof(cplx_t) * nt);
    for (int i = 0; i < nt; ++i) {
        double a = s->COLLISION_TAU * s->OMEGA_d[i];
        double k = creal(s->WAVENUMBER_d[i]);
        double _Complex pf = -(a / (1.0 + I * a) / 2.0 / k * O_ELECTRON_CHARGE * O_ELECTRON_CHARGE / m * O_VACUUM_PERMEABILITY);
        s->PLASMA_FUNC[i] = (cplx_t)pf;
    }
    /* calc_ionization_rate, ionization.cpp:313-331 */
    {
        char ip[512];
        o_load_string(s, "IONIZATION_PROFILE", ip, 512, 1, "NO");
        if (strncmp(ip, "NO", 512) != 0) o_throw(s, "IONIZATION_PROFILE files are not supported by this build");
        for (int i = 0; i < O_ION_N - 1; ++i)
            s->IONIZATION_RATE_LN[i] = o_keldysh_rate_ln(s, O_ION_MIN_LN + O_ION_LN_STEP * i) - O_IONRATE_DENOM_LN;
        s->IONIZATION_RATE_LN[O_ION_N - 1] = 0.0;
    }
    /* calc_raman_function, solver_main.cpp:401-413 */
    {
        double M = s->TSTEP * (1 + s->OMEGA_RAMAN * s->OMEGA_RAMAN * s->TAU_RAMAN * s->TAU_RAMAN) / s->TAU_RAMAN / (s->OMEGA_RAMAN * s->TAU_RAMAN);
        cplx_t* r = s->lb[0].b1;
        s->RAMAN_FUNCTION = (cplx_t*)malloc(sizeof(cplx_t) * nt);
        for (int i = 0; i < nt; ++i) {
            double t = s->TSTEP * i;
            r[i] = (real_t)(s->RAMAN_FRACTION * M * exp(-t / s->TAU_RAMAN) * sin(s->OMEGA_RAMAN * t));
        }
        o_fft_exec(&s->fft, r, s->RAMAN_FUNCTION, -1);
        for (int i = 0; i < nt; ++i) s->RAMAN_FUNCTION[i] += (real_t)(1 - s->RAMAN_FRACTION);
    }
    s->IONIZATION_POTENTIAL *= O_IONRATE_DENOM / O_INTENSITY_DENOM;
    s->NONLIN_REFRINDEX *= O_INTENSITY_DENOM;
    s->RECOMBINATION_TAU *= O_IONRATE_DENOM;
    s->PONDEROMOTIVE_COEFFICIENT *= O_IONRATE_DENOM;

    /* march-precision copies */
    s->OMEGA = (real_t*)malloc(sizeof(real_t) * nt);
    s->WAVENUMBER = (cplx_t*)malloc(sizeof(cplx_t) * nt);
    s->HO = (real_t*)malloc(sizeof(real_t) * nt);
    s->KERR_PROFILE = (real_t*)malloc(sizeof(real_t) * nt);
    for (int i = 0; i < nt; ++i) {
        s->OMEGA[i] = (real_t)s->OMEGA_d[i];
        s->WAVENUMBER[i] = (cplx_t)s->WAVENUMBER_d[i];
        s->HO[i] = (real_t)s->HO_d[i];
    }
    /* load_kerr_th_profiles, process_input.cpp:966-974 */
    {
        char kp[300];
        o_load_string(s, "KERR_PROFILE", kp, 300, 1, "NO");
        if (strncmp(kp, "NO", 300) != 0) o_throw(s, "KERR_PROFILE files are not supported by this build");
        for (int i = 0; i < nt; ++i)
            s->KERR_PROFILE[i] = (real_t)(s->NONLIN_REFRINDEX * s->OMEGA_d[i] / O_LIGHT_VELOCITY / creal(s->WAVENUMBER_d[i]) * s->OMEGA_d[i] / O_LIGHT_VELOCITY);
    }
    s->CURRENT_Z = s->ZNET[0];

    /* ---- create_mystartcondition ---- */
    char shape[200];
    o_load_string(s, "PULSE_SHAPE", shape, 200, 0, "");
    if (strncmp("gg", shape, 2) != 0) o_throw(s, "only PULSE_SHAPE gg is supported by this build");
    create_gg(s);
    for (size_t i = 0; i < n; ++i) s->FIELD[i] /= (real_t)O_FIELD_DENOM;

    /* BIGBUFFER1 = IFFT(FIELD)/N_T; init_zstep_kerr (process_input.cpp:734-753) */
    double maxI = 0;
    for (int r = 0; r < nr; ++r) {
        o_fft_exec(&s->fft, s->FIELD + (size_t)r * nt, s->BIG1 + (size_t)r * nt, +1);
        for (int t = 0; t < nt; ++t) {
            cplx_t v = s->BIG1[(size_t)r * nt + t] / (real_t)nt;
            s->BIG1[(size_t)r * nt + t] = v;
            double a = creal(v) * creal(v) + cimag(v) * cimag(v);
            if (a > maxI) maxI = a;
        }
    }
    s->ZSTEP = s->ZNET[1] - s->ZNET[0];
    double kNL = w0 / O_LIGHT_VELOCITY * s->NONLIN_REFRINDEX * maxI;
    if (s->ZSTEP * kNL > s->MAX_TOLERANCE) s->ZSTEP = s->MAX_TOLERANCE / kNL;

    memcpy(s->BIG2, s->FIELD, sizeof(cplx_t) * n);
    o_qdht_run_many(s, s->BIG2);
}
