"""ctypes front end of the oracle: the CPU restatement of the reference's
gsolver_solid_qdht z-march (see oracle.h for the reference files it follows).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product (gsolver_amd/) never
imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def build(force=False):
    """Compile the oracle with its Makefile (gcc only)."""
    need = force or not all(
        os.path.exists(os.path.join(HERE, f)) for f in ("liboracle_fp64.so", "liboracle_fp32.so"))
    if need:
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)


def lib(precision="fp64"):
    if precision not in _LIBS:
        path = os.path.join(HERE, "liboracle_%s.so" % precision)
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp, cp, dp, ip = C.c_void_p, C.c_char_p, C.c_double, C.c_int
        D = C.POINTER(C.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [cp, ip, C.POINTER(cp), cp, ip, cp, ip]
        L.orc_destroy.argtypes = [vp]
        L.orc_error.restype = cp
        L.orc_error.argtypes = [vp]
        L.orc_dims.argtypes = [vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip)]
        L.orc_scalar.restype = dp
        L.orc_scalar.argtypes = [vp, cp]
        L.orc_set_scalar.argtypes = [vp, cp, dp]
        L.orc_get.argtypes = [vp, cp, D]
        L.orc_put.argtypes = [vp, cp, D]
        L.orc_nl_response.argtypes = [vp, D, D]
        L.orc_plasma.argtypes = [vp, D, D, D]
        L.orc_photoionization.restype = dp
        L.orc_photoionization.argtypes = [vp, dp, dp, dp]
        L.orc_keldysh_rate_ln.restype = dp
        L.orc_keldysh_rate_ln.argtypes = [vp, dp]
        L.orc_qdht_apply.argtypes = [vp, D]
        L.orc_nl_step.argtypes = [vp, D, D]
        L.orc_lstep.argtypes = [vp]
        L.orc_propagator.argtypes = [vp]
        L.orc_fixed_step.argtypes = [vp, dp, D, D]
        L.orc_fixed_step2.argtypes = [vp, dp, dp]
        L.orc_nl_response_vjp.argtypes = [vp, D, D, D]
        L.orc_nl_step_vjp.argtypes = [vp, D, D, dp, D]
        L.orc_adjoint_march.argtypes = [vp, ip, D, D, D, D, D, D]
        L.orc_run.argtypes = [vp, cp, C.c_long]
        L.orc_fft.argtypes = [ip, D, D, ip]
        for f in ("orc_bessel_j0", "orc_bessel_j1", "orc_dawson", "orc_sqrtho"):
            getattr(L, f).restype = dp
            getattr(L, f).argtypes = [dp]
        L.orc_bessel_zero_j0.restype = dp
        L.orc_bessel_zero_j0.argtypes = [ip]
        L.orc_ellipke.argtypes = [dp, D, D]
        L.orc_raman_circular.restype = dp
        L.orc_raman_circular.argtypes = [ip, D, dp, dp, dp, dp, D]
        _LIBS[precision] = L
    return _LIBS[precision]


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def cplx_in(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64), a


class OracleError(RuntimeError):
    pass


def fft(x, sign=-1, precision="fp64"):
    """FFTW-convention DFT of a complex vector (sign -1 forward, +1 backward)."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib(precision).orc_fft(len(x), _dptr(x.view(np.float64)), _dptr(out.view(np.float64)), sign)
    return out


class Sim:
    """One oracle simulation, created from a reference-format config file."""

    def __init__(self, cfg, overrides=(), workdir=None, nthreads=1, precision="fp64"):
        self.L = lib(precision)
        self.precision = precision
        ov = [str(v).encode() for kv in overrides for v in kv] if overrides and isinstance(overrides[0], (tuple, list)) \
            else [str(v).encode() for v in overrides]
        arr = (C.c_char_p * max(1, len(ov)))(*ov)
        err = C.create_string_buffer(512)
        cwd = os.getcwd()
        cfg_dir = os.path.dirname(os.path.abspath(cfg))
        # the reference resolves DISPERSION_FILE and dump names relative to its cwd
        os.chdir(cfg_dir)
        try:
            self.h = self.L.orc_create(os.path.abspath(cfg).encode(), len(ov), arr,
                                       (workdir or cfg_dir).encode(), nthreads, err, 512)
        finally:
            os.chdir(cwd)
        if not self.h:
            raise OracleError(err.value.decode())
        self.cfg_dir = cfg_dir
        R, T, NZ = C.c_int(), C.c_int(), C.c_int()
        self.L.orc_dims(self.h, C.byref(R), C.byref(T), C.byref(NZ))
        self.R, self.T, self.NZ = R.value, T.value, NZ.value

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self):
        raise OracleError(self.L.orc_error(self.h).decode())

    def scalar(self, name):
        return self.L.orc_scalar(self.h, name.encode())

    def set_scalar(self, name, v):
        if self.L.orc_set_scalar(self.h, name.encode(), float(v)):
            raise KeyError(name)

    _SHAPES = {"omega": ("T", False), "wavenum": ("T", True), "ho": ("T", False), "kerr": ("T", False),
               "plasma_func": ("T", True), "raman_function": ("T", True), "ionrate_ln": (8192, False),
               "qdht_x": ("R", False), "qdht_m1": ("R", False), "qdht_c": ("RR", False), "znet": ("NZ", False),
               "field": ("RT", True), "b1": ("RT", True), "b2": ("RT", True)}

    def get(self, name):
        shp, cplx = self._SHAPES[name]
        n = {"T": self.T, "R": self.R, "NZ": self.NZ, "RR": self.R * self.R, "RT": self.R * self.T}.get(shp, shp)
        out = np.zeros(n, dtype=np.complex128 if cplx else np.float64)
        if self.L.orc_get(self.h, name.encode(), _dptr(out.view(np.float64))):
            raise KeyError(name)
        if shp == "RT":
            out = out.reshape(self.R, self.T)
        if shp == "RR":
            out = out.reshape(self.R, self.R).T.copy()   # stored C[i + R*j] -> [i, j]
        return out

    def put(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.complex128).reshape(-1)
        if self.L.orc_put(self.h, name.encode(), _dptr(a.view(np.float64))):
            raise KeyError(name)

    def nl_response(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        out = np.empty_like(a)
        if self.L.orc_nl_response(self.h, _dptr(a.view(np.float64)), _dptr(out.view(np.float64))):
            self._err()
        return out

    def plasma(self, e):
        e = np.ascontiguousarray(e, dtype=np.complex128)
        rho = np.empty(len(e))
        loss = np.empty_like(e)
        if self.L.orc_plasma(self.h, _dptr(e.view(np.float64)), _dptr(rho), _dptr(loss.view(np.float64))):
            self._err()
        return rho, loss

    def photoionization(self, re, im, rho=0.0):
        v = self.L.orc_photoionization(self.h, re, im, rho)
        if np.isnan(v):
            self._err()
        return v

    def keldysh_rate_ln(self, lnI):
        return self.L.orc_keldysh_rate_ln(self.h, lnI)

    def qdht(self, data):
        d = np.ascontiguousarray(data, dtype=np.complex128).copy()
        self.L.orc_qdht_apply(self.h, _dptr(d.view(np.float64)))
        return d

    def nl_step(self):
        a, b = C.c_double(), C.c_double()
        if self.L.orc_nl_step(self.h, C.byref(a), C.byref(b)):
            self._err()
        return a.value, b.value

    def lstep(self):
        self.L.orc_lstep(self.h)

    def propagator(self):
        if self.L.orc_propagator(self.h):
            self._err()

    def fixed_step(self, dz):
        a, b = C.c_double(), C.c_double()
        if self.L.orc_fixed_step(self.h, dz, C.byref(a), C.byref(b)):
            self._err()
        return a.value, b.value

    def fixed_step2(self, dz_rk, dz_lin):
        if self.L.orc_fixed_step2(self.h, dz_rk, dz_lin):
            self._err()

    # ---- reverse mode (o_adjoint.c); cotangent of complex z: dL/dRe z + i dL/dIm z
    def nl_response_vjp(self, a, lam):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        lam = np.ascontiguousarray(lam, dtype=np.complex128)
        out = np.empty_like(a)
        if self.L.orc_nl_response_vjp(self.h, _dptr(a.view(np.float64)), _dptr(lam.view(np.float64)),
                                      _dptr(out.view(np.float64))):
            self._err()
        return out

    def nl_step_vjp(self, A, g, dz):
        A = np.ascontiguousarray(A, dtype=np.complex128)
        g = np.ascontiguousarray(g, dtype=np.complex128)
        out = np.empty_like(A)
        if self.L.orc_nl_step_vjp(self.h, _dptr(A.view(np.float64)), _dptr(g.view(np.float64)), float(dz),
                                  _dptr(out.view(np.float64))):
            self._err()
        return out

    def adjoint_march(self, dz_rk, dz_lin, gfield, gb2=None):
        """March len(dz_rk) fixed steps from the current state, then return the
        cotangents (gfield0, gb2_0) of the initial FIELD and BIG2."""
        dz_rk = np.ascontiguousarray(dz_rk, dtype=np.float64)
        dz_lin = np.ascontiguousarray(dz_lin, dtype=np.float64)
        gf = np.ascontiguousarray(gfield, dtype=np.complex128).reshape(self.R, self.T)
        g0 = np.empty_like(gf)
        b0 = np.empty_like(gf)
        gbp = None
        if gb2 is not None:
            gb = np.ascontiguousarray(gb2, dtype=np.complex128).reshape(self.R, self.T)
            gbp = _dptr(gb.view(np.float64))
        if self.L.orc_adjoint_march(self.h, len(dz_rk), _dptr(dz_rk), _dptr(dz_lin), _dptr(gf.view(np.float64)),
                                    gbp, _dptr(g0.view(np.float64)), _dptr(b0.view(np.float64))):
            self._err()
        return g0, b0

    def run(self, outdir, max_steps=0):
        cwd = os.getcwd()
        os.chdir(self.cfg_dir)
        try:
            rc = self.L.orc_run(self.h, os.path.abspath(outdir).encode(), max_steps)
        finally:
            os.chdir(cwd)
        if rc < 0:
            self._err()
        return rc


def load_udf(path):
    """Read a .udf dump (dump_types.cpp:25-72): 'UDF1.0\\0', 6 int32, slabs."""
    raw = open(path, "rb").read()
    if raw[:7] != b"UDF1.0\0":
        raise ValueError("not a UDF1.0 file")
    hdr = np.frombuffer(raw[7:31], dtype=np.int32).copy()
    fs, cplx, tN, xN, yN, zN = (int(v) for v in hdr)
    if fs == 8:
        dt = np.complex128 if cplx else np.float64
    else:
        dt = np.complex64 if cplx else np.float32
    data = np.frombuffer(raw[31:], dtype=dt)
    n = yN * xN * tN
    return hdr, data[: (len(data) // n) * n].reshape(-1, n)
