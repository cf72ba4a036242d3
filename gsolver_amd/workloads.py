"""Reference-format inputs for the benchmark and test workloads.

- write_sellmeier_rawdisp(): a type-3 .rawdisp table (the layout of
  process_input.cpp:687-693 / m/refractive_index_raw_write.m:5-16:
  int32 3, int32 N, double omega[N], double n[2N] as (re, im)) from Sellmeier
  coefficients, since the reference's HO-dispersion step only reads raw tables
  (process_input.cpp:864).
- write_config(): a whitespace "NAME VALUE" config for the named workloads:
  "c0"  256 x 1024, 0.5 mm / 400 fs, 2 cm, 1 uJ 50 fs, w0 = 60 um, fused silica
  "c1"  1024 x 4096, 1 mm / 800 fs, 3 cm, 5 uJ 40 fs, w0 = 100 um, 1 % seeded
        amplitude noise (NOISE_MAP_LEVEL / NOISE_SEED)
  "c2"  c1 with dumps at 0, 1, 2, 3 cm
  "c3"  the c0 medium and pulse on 512 x 2048 over 1 cm (the reverse-mode
        workload; the pulse energy is a keyword, see write_config)
- batch_configs(): the energy / width scan of 64 seeded runs over uneven grids.
"""
import os

import numpy as np

C_LIGHT = 299792458.0
# fused silica, 3-term Sellmeier (B_i, C_i in um^2)
SIO2_B = (0.6961663, 0.4079426, 0.8974794)
SIO2_C = (0.0684043 ** 2, 0.1162414 ** 2, 9.896161 ** 2)


def sellmeier_n(lam_um, B=SIO2_B, Cc=SIO2_C):
    l2 = np.asarray(lam_um, dtype=float) ** 2
    n2 = 1.0 + sum(b * l2 / (l2 - c) for b, c in zip(B, Cc))
    return np.sqrt(n2)


def write_sellmeier_rawdisp(path, lam_min_um=0.25, lam_max_um=4.0, npts=20001, B=SIO2_B, Cc=SIO2_C):
    lam = np.linspace(lam_max_um, lam_min_um, npts)        # ascending omega
    omega = 2 * np.pi * C_LIGHT / (lam * 1e-6)
    n = sellmeier_n(lam, B, Cc)
    with open(path, "wb") as f:
        np.array([3, npts], dtype=np.int32).tofile(f)
        omega.astype(np.float64).tofile(f)
        np.stack([n, np.zeros_like(n)], axis=1).astype(np.float64).tofile(f)
    return path


def workload_keys(name):
    common = dict(LAMBDA_V="0.8e-6", DISPERSION_FILE="fused_silica_sellmeier.rawdisp", ZNET_TYPE="e",
                  PULSE_SHAPE="gg", WAVEFRONT_RADIUS="1e30", NONLIN_REFRINDEX="2.5e-20", RAMAN_FRACTION="0.18",
                  OMEGA_RAMAN="%.6e" % (2 * np.pi * 13.2e12), TAU_RAMAN="32e-15", NEUTRAL_DENSITY="2.1e28",
                  RECOMBINATION_TAU="150e-15", COLLISION_TAU="23e-15", IONIZATION_POTENTIAL="9",
                  AMBIENT_CARRIERS="0", MAX_TOLERANCE="1e-4", DUMPS_N="0")
    fw = np.sqrt(2 * np.log(2))   # DIAMETER_FWHM = w0 sqrt(2 ln 2) (pulse_gg.cpp:40-43)
    if name == "c0":
        return dict(common, N_T="1024", T_MIN="-200e-15", T_MAX="200e-15", N_R="256", R_MAX="0.5e-3", N_Z="3",
                    Z_MIN="0", Z_MAX="2e-2", DURATION_FWHM="50e-15", DIAMETER_FWHM="%.6e" % (60e-6 * fw),
                    ENERGY="1e-6")
    if name == "c3":
        return dict(workload_keys("c0"), N_R="512", N_T="2048", N_Z="2", Z_MAX="1e-2")
    if name in ("c1", "c2"):
        k = dict(common, N_T="4096", T_MIN="-400e-15", T_MAX="400e-15", N_R="1024", R_MAX="1e-3", N_Z="4",
                 Z_MIN="0", Z_MAX="3e-2", DURATION_FWHM="40e-15", DIAMETER_FWHM="%.6e" % (100e-6 * fw),
                 ENERGY="5e-6", NOISE_MAP_LEVEL="0.01", NOISE_SEED="12345")
        if name == "c2":
            k.update(DUMPS_N="3", DUMPTYPE0="flux", DUMP_FILENAME0="c2_%s_f.udf", DUMPTYPE1="plasma_max",
                     DUMP_FILENAME1="c2_%s_pm.udf", DUMPTYPE2="field_axis", DUMP_FILENAME2="c2_%s_fa.udf")
        return k
    raise KeyError(name)


def write_config(directory, name, out_name=None, **overrides):
    """Write <directory>/<out_name or name>.txt for workload `name` (keys
    overridden by the keyword arguments) and the fused-silica .rawdisp beside it."""
    os.makedirs(directory, exist_ok=True)
    rd = os.path.join(directory, "fused_silica_sellmeier.rawdisp")
    if not os.path.exists(rd):
        write_sellmeier_rawdisp(rd)
    keys = workload_keys(name)
    keys.update({k: str(v) for k, v in overrides.items()})
    path = os.path.join(directory, (out_name or name) + ".txt")
    with open(path, "w") as f:
        for k, v in keys.items():
            f.write("%s %s\n" % (k, v))
    return path


def batch_configs(directory, nruns=64, z_max=1e-2, seed=20261202, **overrides):
    """The energy / width scan: nruns runs of the c0 medium and pulse with the
    energy log-uniform in 0.5-10 uJ, w0 uniform in 40-150 um, N_R drawn from
    {128, 256, 512} and N_T from {512, 1024, 2048}, all from one seed, each
    marched to z_max (planes at 0 and z_max).  Returns the config paths."""
    rng = np.random.default_rng(seed)
    os.makedirs(directory, exist_ok=True)
    out = []
    for i in range(nruns):
        e = float(np.exp(rng.uniform(np.log(0.5e-6), np.log(10e-6))))
        w0 = float(rng.uniform(40e-6, 150e-6))
        nr = int(rng.choice([128, 256, 512]))
        nt = int(rng.choice([512, 1024, 2048]))
        keys = dict(N_R=nr, N_T=nt, ENERGY="%.6e" % e, DIAMETER_FWHM="%.6e" % (w0 * np.sqrt(2 * np.log(2))),
                    N_Z=2, Z_MAX=z_max)
        keys.update(overrides)
        out.append(write_config(directory, "c0", "run%02d" % i, **keys))
    return out
