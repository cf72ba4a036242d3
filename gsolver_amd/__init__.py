"""gsolver_amd -- MI355X (gfx950) implementation of the gsolver_solid_qdht
per-z-step UPPE right-hand side and RK march.

This module is a thin ctypes front end over the C-ABI of
include/gsolver_amd.h (libgsolver_amd.so, built in-tree by
gsolver_amd/build.py).  All array work runs in the HIP kernels of
gsolver_amd/csrc; there is no CPU fallback: loading fails loudly when the
library is missing, and creating a plan fails without a GPU.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgsolver_amd.so")

FP64, FP32, MIXED = 0, 1, 2
PRECISIONS = {"fp64": FP64, "fp32": FP32, "mixed": MIXED}
FIELD, B1, B2, FK = 0, 1, 2, 3

_lib = None


class GsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code
        self.msg = msg


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgsolver_amd.so is not built: run `python -m gsolver_amd.build` "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp, cp, ip, dp, sz = C.c_void_p, C.c_char_p, C.c_int, C.c_double, C.c_size_t
        D = C.POINTER(C.c_double)
        L.gs_plan_create.argtypes = [cp, ip, C.POINTER(cp), ip, cp, C.POINTER(vp)]
        L.gs_plan_destroy.argtypes = [vp]
        L.gs_last_error.restype = cp
        L.gs_setup_table.argtypes = [cp, ip, C.POINTER(cp), cp, D, sz]
        L.gs_dims.argtypes = [vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip)]
        L.gs_get_scalar.argtypes = [vp, cp, D]
        L.gs_set_scalar.argtypes = [vp, cp, dp]
        L.gs_get_table.argtypes = [vp, cp, D, sz]
        L.gs_get.argtypes = [vp, ip, D]
        L.gs_put.argtypes = [vp, ip, D]
        L.gs_nl_step.argtypes = [vp, dp, D, D]
        L.gs_nl_response.argtypes = [vp, D, D, ip]
        L.gs_nl_response_noalloc.argtypes = [vp, D, D, ip]
        L.gs_qdht.argtypes = [vp, ip]
        L.gs_linear_step.argtypes = [vp, dp]
        L.gs_step.argtypes = [vp]
        L.gs_fixed_step.argtypes = [vp, dp, D, D]
        L.gs_march.argtypes = [vp, cp, C.c_long]
        L.gs_line_diagnostics.argtypes = [vp, D, D, D]
        L.gs_fft_lines.argtypes = [vp, D, D, ip, ip]
        L.gs_last_timings.argtypes = [vp, D]
        L.gs_fixed_steps.argtypes = [vp, dp, ip, D]
        L.gs_march_many.argtypes = [C.POINTER(vp), ip, C.POINTER(cp), C.POINTER(ip)]
        L.gs_adjoint_forward.argtypes = [vp, ip, D, D]
        L.gs_adjoint_backward.argtypes = [vp, D, D, ip, D, D]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc):
    if rc < 0:
        raise GsError(rc, lib().gs_last_error().decode())
    return rc


def setup_table(cfg, name, n, cplx=False, overrides=()):
    """Host set-up only (no GPU): the table `name` gs_plan_create would build."""
    ov = [str(v).encode() for v in overrides]
    arr = (C.c_char_p * max(1, len(ov)))(*ov)
    out = np.zeros(n, dtype=np.complex128 if cplx else np.float64)
    cfg = os.path.abspath(cfg)
    cwd = os.getcwd()
    os.chdir(os.path.dirname(cfg))
    try:
        _check(lib().gs_setup_table(cfg.encode(), len(ov), arr, name.encode(), _d(out.view(np.float64)),
                                    out.view(np.float64).size))
    finally:
        os.chdir(cwd)
    return out


def march_many(plans, outdirs):
    """gs_march of several plans at once (their steps overlap on the device)."""
    n = len(plans)
    hs = (C.c_void_p * n)(*[p.h for p in plans])
    ds = (C.c_char_p * n)(*[os.path.abspath(d).encode() for d in outdirs])
    rcs = (C.c_int * n)()
    cwd = os.getcwd()
    os.chdir(plans[0].cfg_dir)
    try:
        _check(lib().gs_march_many(hs, n, ds, rcs))
    finally:
        os.chdir(cwd)
    return list(rcs)


class Plan:
    """One run of the z-march on the GPU, set up from a reference config file
    (process_input semantics, relative paths resolved against the config's
    directory)."""

    def __init__(self, cfg, overrides=(), precision="mixed", workdir=None):
        L = lib()
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        ov = [str(v).encode() for v in overrides]
        arr = (C.c_char_p * max(1, len(ov)))(*ov)
        h = C.c_void_p()
        cfg = os.path.abspath(cfg)
        cwd = os.getcwd()
        os.chdir(os.path.dirname(cfg))
        try:
            _check(L.gs_plan_create(cfg.encode(), len(ov), arr, prec,
                                    workdir.encode() if workdir else None, C.byref(h)))
        finally:
            os.chdir(cwd)
        self.h = h
        self.cfg_dir = os.path.dirname(cfg)
        self.precision = prec
        R, T, NZ = C.c_int(), C.c_int(), C.c_int()
        L.gs_dims(h, C.byref(R), C.byref(T), C.byref(NZ))
        self.R, self.T, self.NZ = R.value, T.value, NZ.value

    def close(self):
        if getattr(self, "h", None):
            lib().gs_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def scalar(self, name):
        v = C.c_double()
        _check(lib().gs_get_scalar(self.h, name.encode(), C.byref(v)))
        return v.value

    def set_scalar(self, name, v):
        _check(lib().gs_set_scalar(self.h, name.encode(), float(v)))

    _TABLES = {"omega": ("T", False), "wavenum": ("T", True), "ho": ("T", False), "kerr": ("T", False),
               "plasma_func": ("T", True), "raman_function": ("T", True), "ionrate_ln": (8192, False),
               "qdht_x": ("R", False), "qdht_m1": ("R", False), "qdht_c": ("RR", False), "znet": ("NZ", False)}

    def table(self, name):
        shp, cplx = self._TABLES[name]
        n = {"T": self.T, "R": self.R, "NZ": self.NZ, "RR": self.R * self.R}.get(shp, shp)
        out = np.zeros(n, dtype=np.complex128 if cplx else np.float64)
        _check(lib().gs_get_table(self.h, name.encode(), _d(out.view(np.float64)), out.view(np.float64).size))
        return out.reshape(self.R, self.R) if shp == "RR" else out

    def get(self, buf=FIELD):
        out = np.empty((self.R, self.T), dtype=np.complex128)
        _check(lib().gs_get(self.h, buf, _d(out.view(np.float64))))
        return out

    def put(self, arr, buf=FIELD):
        a = np.ascontiguousarray(arr, dtype=np.complex128).reshape(self.R, self.T)
        _check(lib().gs_put(self.h, buf, _d(a.view(np.float64))))

    def nl_step(self, dz):
        a, b = C.c_double(), C.c_double()
        _check(lib().gs_nl_step(self.h, dz, C.byref(a), C.byref(b)))
        return a.value, b.value

    def nl_response(self, lines):
        x = np.ascontiguousarray(np.atleast_2d(lines), dtype=np.complex128)
        out = np.empty_like(x)
        _check(lib().gs_nl_response(self.h, _d(x.view(np.float64)), _d(out.view(np.float64)), x.shape[0]))
        return out

    def nl_response_noalloc(self, lines):
        """nl_response through the allocation-free entry point (fixed device scratch)."""
        x = np.ascontiguousarray(np.atleast_2d(lines), dtype=np.complex128)
        out = np.empty_like(x)
        _check(lib().gs_nl_response_noalloc(self.h, _d(x.view(np.float64)), _d(out.view(np.float64)), x.shape[0]))
        return out

    def qdht(self, buf):
        _check(lib().gs_qdht(self.h, buf))

    def linear_step(self, dz):
        _check(lib().gs_linear_step(self.h, dz))

    def step(self):
        _check(lib().gs_step(self.h))

    def fixed_step(self, dz):
        a, b = C.c_double(), C.c_double()
        _check(lib().gs_fixed_step(self.h, dz, C.byref(a), C.byref(b)))
        return a.value, b.value

    def fixed_steps(self, dz, n, timings=True):
        """n fixed-dz steps with one synchronisation; returns the mean kernel
        times per step (timings=False: no timing events between the kernels,
        returns None)."""
        if not timings:
            _check(lib().gs_fixed_steps(self.h, dz, int(n), None))
            return None
        t = np.zeros(4)
        _check(lib().gs_fixed_steps(self.h, dz, int(n), _d(t)))
        return dict(zip(("nl_rk4", "qdht_lstep", "qdht_back", "reduce"), t))

    def march(self, outdir, max_steps=0):
        cwd = os.getcwd()
        os.chdir(self.cfg_dir)
        try:
            return _check(lib().gs_march(self.h, os.path.abspath(outdir).encode(), max_steps))
        finally:
            os.chdir(cwd)

    def line_diagnostics(self):
        f, m, r = (np.empty(self.R) for _ in range(3))
        _check(lib().gs_line_diagnostics(self.h, _d(f), _d(m), _d(r)))
        return f, m, r

    def fft_lines(self, lines, sign):
        x = np.ascontiguousarray(np.atleast_2d(lines), dtype=np.complex128)
        out = np.empty_like(x)
        _check(lib().gs_fft_lines(self.h, _d(x.view(np.float64)), _d(out.view(np.float64)), x.shape[0], sign))
        return out

    # ---- reverse mode: cotangent of complex z is dL/dRe z + i dL/dIm z
    def adjoint_forward(self, dz_rk, dz_lin):
        """March len(dz_rk) fixed steps (NL RK4 at dz_rk[k], linear step at
        dz_lin[k]) keeping FIELD before every step on the device."""
        a = np.ascontiguousarray(dz_rk, dtype=np.float64)
        b = np.ascontiguousarray(dz_lin, dtype=np.float64)
        if a.shape != b.shape:
            raise ValueError("dz_rk and dz_lin differ in length")
        _check(lib().gs_adjoint_forward(self.h, len(a), _d(a), _d(b)))

    def adjoint_backward(self, gfield, gb2=None, tie_b2=True):
        """Cotangents (gfield0, gb2_0) of the initial FIELD and BIGBUFFER2 of the
        recorded march from those of the final ones; with tie_b2 the BIGBUFFER2
        part is also folded into gfield0 (initial BIGBUFFER2 = QDHT(FIELD))."""
        g = np.ascontiguousarray(gfield, dtype=np.complex128).reshape(self.R, self.T)
        gbp = None
        if gb2 is not None:
            gb = np.ascontiguousarray(gb2, dtype=np.complex128).reshape(self.R, self.T)
            gbp = _d(gb.view(np.float64))
        g0 = np.empty((self.R, self.T), dtype=np.complex128)
        b0 = np.empty_like(g0)
        _check(lib().gs_adjoint_backward(self.h, _d(g.view(np.float64)), gbp, 1 if tie_b2 else 0,
                                         _d(g0.view(np.float64)), _d(b0.view(np.float64))))
        return g0, b0

    def timings(self):
        t = np.zeros(4)
        _check(lib().gs_last_timings(self.h, _d(t)))
        return {"nl_rk4": t[0], "qdht_lstep": t[1], "qdht_back": t[2], "reduce": t[3]}
