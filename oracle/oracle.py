"""ctypes wrapper of the FP64 CPU oracle (oracle/dqmc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / CPU baseline, never by the product.
Matrices cross the boundary column-major (Fortran order), as the reference's Eigen
matrices are (include/dqmc_core.h:27).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)
_llp = C.POINTER(C.c_longlong)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        sig = {
            "orc_set_mkl": (C.c_int, [C.c_int]),
            "orc_rng_set_fma": (None, [C.c_int]),
            "orc_rng_stream": (None, [C.c_uint64, C.c_int, _u64p]),
            "orc_rng_canonical_stream": (None, [C.c_uint64, C.c_int, _dp]),
            "orc_nn_hoppings": (None, [C.c_int, _dp]),
            "orc_expk": (None, [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp]),
            "orc_fft_eigens": (None, [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp]),
            "orc_set_fft": (None, [C.c_void_p, C.c_int]),
            "orc_svd": (C.c_int, [C.c_int, _dp, _dp, _dp, _dp]),
            "orc_inverse": (C.c_int, [C.c_int, _dp, _dp]),
            "orc_fullpiv_qr_solve": (None, [C.c_int, _dp, _dp, _dp]),
            "orc_create": (vp, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_int, C.c_int, C.c_int]),
            "orc_destroy": (None, [vp]),
            "orc_set_seed": (None, [vp, C.c_uint64]),
            "orc_get_rng_state": (C.c_uint64, [vp]),
            "orc_set_rng_state": (None, [vp, C.c_uint64]),
            "orc_set_thermalization": (None, [vp, C.c_int]),
            "orc_alpha": (C.c_double, [vp]),
            "orc_fields_random": (None, [vp]),
            "orc_set_fields": (None, [vp, _ip]),
            "orc_get_fields": (None, [vp, _ip]),
            "orc_get_expk": (None, [vp, _dp, _dp]),
            "orc_apply_bop": (None, [vp, C.c_int, C.c_int, _dp]),
            "orc_init": (C.c_int, [vp]),
            "orc_sweep_up": (C.c_int, [vp]),
            "orc_sweep_down": (C.c_int, [vp]),
            "orc_sweep_dynamic": (C.c_int, [vp]),
            "orc_get_gf": (C.c_int, [vp, C.c_int, C.c_int, _dp]),
            "orc_set_gftt": (None, [vp, _dp]),
            "orc_compare_error": (C.c_double, [C.c_int, _dp, _dp]),
            "orc_acceptance_ratio": (C.c_double, [vp, C.c_int, C.c_int]),
            "orc_update_green_function": (None, [vp, C.c_int, C.c_int]),
            "orc_flip_field": (None, [vp, C.c_int, C.c_int]),
            "orc_get_stats": (None, [vp, _dp, _dp, _llp, _llp]),
            "orc_get_accmask": (None, [vp, _u32p]),
            "orc_current_t": (C.c_int, [vp]),
            "orc_equaltime_gf_at": (C.c_int, [vp, C.c_int, _dp]),
            "orc_measure_equaltime": (None, [vp, _dp, _dp, _dp]),
            "orc_measure_dynamic": (None, [vp, C.c_int, _dp, _dp, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def set_mkl(on):
    return lib().orc_set_mkl(1 if on else 0)


def rng_raw(seed, n):
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_rng_stream(seed, n, _ptr(out, _u64p))
    return out


def rng_canonical(seed, n, fma=True):
    out = np.zeros(n)
    lib().orc_rng_set_fma(1 if fma else 0)
    try:
        lib().orc_rng_canonical_stream(seed, n, _ptr(out))
    finally:
        lib().orc_rng_set_fma(1)
    return out


def nn_hoppings(nl):
    n = nl * nl
    a = np.zeros((n, n), order="F")
    lib().orc_nn_hoppings(nl, _ptr(a))
    return a


def expk(nl, dt, nnt, mu):
    n = nl * nl
    e = np.zeros((n, n), order="F")
    ie = np.zeros((n, n), order="F")
    lib().orc_expk(nl, dt, nnt, mu, _ptr(e), _ptr(ie))
    return e, ie


def fft_eigens(nl, dt, nnt, mu):
    """e^{-dt E(k)} and e^{+dt E(k)} at k = kx + nl*ky: the diagonal of (inv)expK in momentum space."""
    n = nl * nl
    lf, li = np.zeros(n), np.zeros(n)
    lib().orc_fft_eigens(nl, dt, nnt, mu, _ptr(lf), _ptr(li))
    return lf, li


def svd(a):
    a = np.asfortranarray(a, dtype=np.float64)
    n = a.shape[0]
    u = np.zeros((n, n), order="F")
    v = np.zeros((n, n), order="F")
    s = np.zeros(n)
    rc = lib().orc_svd(n, _ptr(a), _ptr(u), _ptr(s), _ptr(v))
    return rc, u, s, v


def inverse(a):
    a = np.asfortranarray(a, dtype=np.float64)
    out = np.zeros_like(a, order="F")
    rc = lib().orc_inverse(a.shape[0], _ptr(a), _ptr(out))
    return rc, out


def compare_error(a, b):
    """StableNumerics::matrix_compare_error: the signed max of (a - b)."""
    a = np.asfortranarray(a, dtype=np.float64)
    b = np.asfortranarray(b, dtype=np.float64)
    return float(lib().orc_compare_error(a.shape[0], _ptr(a), _ptr(b)))


def fullpiv_qr_solve(h, b):
    h = np.asfortranarray(h, dtype=np.float64)
    b = np.asfortranarray(b, dtype=np.float64)
    x = np.zeros_like(b, order="F")
    lib().orc_fullpiv_qr_solve(h.shape[0], _ptr(h), _ptr(b), _ptr(x))
    return x


class Chain:
    """One reference Markov chain: DQMC::Core + Model::HubbardAttractiveU restated (FP64, CPU)."""

    def __init__(self, nl, nt, beta, u, mu, nnt=1.0, pace=10, eqtime=False, dynamic=False, is_fft=False):
        self.nl, self.n, self.nt = nl, nl * nl, nt
        self.beta, self.u, self.mu, self.nnt, self.pace = beta, u, mu, nnt, pace
        self._c = lib().orc_create(nl, nt, beta, nnt, u, mu, pace, int(eqtime), int(dynamic))
        if is_fft:
            lib().orc_set_fft(self._c, 1)

    def close(self):
        if self._c:
            lib().orc_destroy(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def alpha(self):
        return lib().orc_alpha(self._c)

    def set_seed(self, seed):
        lib().orc_set_seed(self._c, seed)

    @property
    def rng_state(self):
        return int(lib().orc_get_rng_state(self._c))

    @rng_state.setter
    def rng_state(self, x):
        lib().orc_set_rng_state(self._c, int(x))

    def set_thermalization(self, on):
        lib().orc_set_thermalization(self._c, int(on))

    def fields_random(self):
        lib().orc_fields_random(self._c)

    @property
    def fields(self):
        s = np.zeros((self.nt, self.n), dtype=np.int32)
        lib().orc_get_fields(self._c, _ptr(s, _ip))
        return s

    @fields.setter
    def fields(self, s):
        s = np.ascontiguousarray(s, dtype=np.int32).reshape(self.nt, self.n)
        lib().orc_set_fields(self._c, _ptr(s, _ip))

    def expk(self):
        e = np.zeros((self.n, self.n), order="F")
        ie = np.zeros((self.n, self.n), order="F")
        lib().orc_get_expk(self._c, _ptr(e), _ptr(ie))
        return e, ie

    def apply_bop(self, op, t, g):
        g = np.array(g, dtype=np.float64, order="F", copy=True)
        lib().orc_apply_bop(self._c, op, t, _ptr(g))
        return g

    def init(self):
        return lib().orc_init(self._c)

    def sweep_up(self):
        return lib().orc_sweep_up(self._c)

    def sweep_down(self):
        return lib().orc_sweep_down(self._c)

    def sweep_dynamic(self):
        return lib().orc_sweep_dynamic(self._c)

    def gf(self, kind=0, t=0):
        out = np.zeros((self.n, self.n), order="F")
        rc = lib().orc_get_gf(self._c, kind, t, _ptr(out))
        if rc:
            raise RuntimeError("gf kind %d not collected" % kind)
        return out

    def acceptance_ratio(self, t, i):
        return lib().orc_acceptance_ratio(self._c, t, i)

    def update_green_function(self, t, i):
        lib().orc_update_green_function(self._c, t, i)

    def flip_field(self, t, i):
        lib().orc_flip_field(self._c, t, i)

    def set_gftt(self, g):
        g = np.asfortranarray(g, dtype=np.float64)
        lib().orc_set_gftt(self._c, _ptr(g))

    def stats(self):
        we, wd = C.c_double(), C.c_double()
        a, v = C.c_longlong(), C.c_longlong()
        lib().orc_get_stats(self._c, C.byref(we), C.byref(wd), C.byref(a), C.byref(v))
        return {"wrap_eq": we.value, "wrap_dyn": wd.value, "accepts": a.value, "visits": v.value}

    def accmask(self):
        out = np.zeros((self.nt, (self.n + 31) // 32), dtype=np.uint32)
        lib().orc_get_accmask(self._c, _ptr(out, _u32p))
        return out

    @property
    def current_t(self):
        return lib().orc_current_t(self._c)

    def equaltime_gf_at(self, t):
        out = np.zeros((self.n, self.n), order="F")
        rc = lib().orc_equaltime_gf_at(self._c, t, _ptr(out))
        return rc, out

    def measure_equaltime(self):
        f, d, s = C.c_double(), C.c_double(), C.c_double()
        lib().orc_measure_equaltime(self._c, C.byref(f), C.byref(d), C.byref(s))
        return f.value, d.value, s.value

    def measure_dynamic(self, kxy):
        kxy = np.ascontiguousarray(kxy, dtype=np.float64).reshape(-1, 2)
        nk = kxy.shape[0]
        gf = np.zeros((nk, self.nt))
        ldos = np.zeros(self.nt)
        spin = np.zeros(self.nt)
        lib().orc_measure_dynamic(self._c, nk, _ptr(kxy), _ptr(gf), _ptr(ldos), _ptr(spin))
        return gf, ldos, spin
