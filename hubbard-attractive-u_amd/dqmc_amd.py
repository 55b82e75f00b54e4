"""Python binding of the MI355X DQMC sweep engine (libdqmc_amd.so, C ABI in include/dqmc_amd.h).

`Engine` mirrors the sweep-path interface of the reference's DQMC::Core and
Model::HubbardAttractiveU (include/dqmc_core.h:76-98, include/hubbard.h:58-73) for a batch
of independent chains.  The shared library must have been built (see Makefile / build());
loading fails loudly when it is missing - there is no CPU fallback.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdqmc_amd.so")

FP64, MIXED = 0, 1
TILE_NAMES = ["64x64", "32x32", "64x32", "64x64pf", "128x128"]   # GEMM tiles of dqmc_get/set_tuning
GF_CURRENT, GF_TT, GF_T0, GF_0T = 0, 1, 2, 3
HEALTH_NONFINITE_G, HEALTH_ZERO_PIVOT, HEALTH_RATIO, HEALTH_WATCHDOG = 1, 2, 4, 8
OBS_FILLING, OBS_DOUBLE_OCCU, OBS_SWAVE, OBS_GREEN_KT, OBS_LDOS, OBS_SPIN = 1, 2, 4, 8, 16, 32
_ERRS = {-1: "EINVAL", -2: "EHIP", -3: "ESTATE", -4: "ENUMERIC", -5: "ENOMEM", -6: "ENODEV"}


class DqmcError(RuntimeError):
    """A failed ABI call; `code` is the DQMC_E* return value."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class Params(C.Structure):
    """dqmc_params_c: the config.toml keys of the sweep path (include/dqmc_params.hpp:20-45)."""
    _fields_ = [("nl", C.c_int), ("nt", C.c_int), ("beta", C.c_double), ("nnt", C.c_double),
                ("u", C.c_double), ("mu", C.c_double), ("stabilization_pace", C.c_int),
                ("is_fft", C.c_int), ("collect_eqtime", C.c_int), ("collect_dynamic", C.c_int)]


_lib = None


def build(force=False):
    args = ["make", "-C", _HERE]
    if force:
        args.append("-B")
    subprocess.check_call(args)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DqmcError("libdqmc_amd.so is not built (%s); run `make -C %s`" % (LIB_PATH, _HERE))
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        sig = {
            "dqmc_create": (C.c_int, [C.POINTER(Params), C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
            "dqmc_destroy": (None, [vp]),
            "dqmc_last_error": (C.c_char_p, [vp]),
            "dqmc_set_chain_mu": (C.c_int, [vp, C.c_int, C.c_double]),
            "dqmc_set_collect": (C.c_int, [vp, C.c_int, C.c_int]),
            "dqmc_set_seed": (C.c_int, [vp, C.c_int, C.c_uint64]),
            "dqmc_get_rng_state": (C.c_int, [vp, C.c_int, C.POINTER(C.c_uint64)]),
            "dqmc_fields_random": (C.c_int, [vp, C.c_int]),
            "dqmc_set_fields": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int8)]),
            "dqmc_get_fields": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int8)]),
            "dqmc_init": (C.c_int, [vp]),
            "dqmc_apply_bop": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp]),
            "dqmc_set_thermalization": (C.c_int, [vp, C.c_int]),
            "dqmc_sweep_up": (C.c_int, [vp]),
            "dqmc_sweep_down": (C.c_int, [vp]),
            "dqmc_sweep_dynamic": (C.c_int, [vp]),
            "dqmc_get_gf": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp]),
            "dqmc_set_gf": (C.c_int, [vp, C.c_int, dp]),
            "dqmc_get_stats": (C.c_int, [vp, C.c_int, dp, dp, dp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
            "dqmc_get_accept_mask": (C.c_int, [vp, C.c_int, C.POINTER(C.c_uint32)]),
            "dqmc_synchronize": (C.c_int, [vp]),
            "dqmc_event_record": (C.c_int, [vp, C.c_int]),
            "dqmc_event_elapsed": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_float)]),
            "dqmc_use_graphs": (C.c_int, [vp, C.c_int]),
            "dqmc_kernel_timer_begin": (C.c_int, [vp, C.c_int, C.c_int]),
            "dqmc_kernel_timer_end": (C.c_int, [vp, dp, C.POINTER(C.c_int)]),
            "dqmc_get_tuning": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float)]),
            "dqmc_set_tuning": (C.c_int, [vp, C.c_int, C.c_int]),
            "dqmc_forward_fixed": (C.c_int, [vp, dp, dp]),
            "dqmc_adjoint": (C.c_int, [vp, dp, dp, dp]),
            "dqmc_dims": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "dqmc_apply_bop_noalloc": (C.c_int, [C.POINTER(Params), C.c_int, C.c_int, vp, vp, vp]),
            "dqmc_measure_setup": (C.c_int, [vp, C.c_uint, C.c_int, C.c_int, dp]),
            "dqmc_measure_equaltime": (C.c_int, [vp]),
            "dqmc_measure_dynamic": (C.c_int, [vp]),
            "dqmc_measure_push_bin": (C.c_int, [vp, C.c_int]),
            "dqmc_get_bins": (C.c_int, [vp, C.c_int, C.c_uint, dp]),
            "dqmc_acceptance_ratio": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp]),
            "dqmc_update_green_function": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
            "dqmc_flip_field": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
            "dqmc_set_watchdog": (C.c_int, [vp, C.c_double]),
            "dqmc_get_health": (C.c_int, [vp, C.c_int, C.POINTER(C.c_uint32)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def apply_bop_noalloc(nl, nt, beta, u, mu, nnt, n_chains, op, d_fields, d_inout, stream=None):
    """dqmc_apply_bop_noalloc on device memory the caller owns (addresses as ints): queues
    operation `op` of Engine.apply_bop on `stream` for n_chains column-major N x N matrices at
    d_inout, with one slice of fields (int8, [chain][N]) at d_fields.  Allocates nothing."""
    p = Params(nl, nt, beta, nnt, u, mu, 1, 0, 0, 0)
    rc = lib().dqmc_apply_bop_noalloc(C.byref(p), n_chains, op, C.c_void_p(d_fields), C.c_void_p(d_inout),
                                      C.c_void_p(stream) if stream else None)
    if rc:
        raise DqmcError("dqmc_apply_bop_noalloc failed: %s %s" % (_ERRS.get(rc, rc), lib().dqmc_last_error(None).decode()),
                        rc)


class Engine:
    """A batch of `n_chains` DQMC chains on one GPU."""

    def __init__(self, nl, nt, beta, u, mu=0.0, nnt=1.0, pace=10, n_chains=1, precision=FP64, delay_k=1,
                 eqtime=False, dynamic=False, is_fft=False):
        self.nl, self.n, self.nt, self.n_chains = nl, nl * nl, nt, n_chains
        p = Params(nl, nt, beta, nnt, u, mu, pace, int(is_fft), int(eqtime), int(dynamic))
        h = C.c_void_p()
        rc = lib().dqmc_create(C.byref(p), n_chains, precision, delay_k, C.byref(h))
        self._h = h
        if rc:
            msg = lib().dqmc_last_error(h if h.value else None)
            if h.value:
                lib().dqmc_destroy(h)
            self._h = C.c_void_p()
            raise DqmcError("dqmc_create failed: %s %s" % (_ERRS.get(rc, rc), (msg or b"").decode()), rc)

    def _chk(self, rc, what):
        if rc:
            raise DqmcError("%s failed: %s %s" % (what, _ERRS.get(rc, rc), lib().dqmc_last_error(self._h).decode()),
                            rc)

    def close(self):
        if self._h and self._h.value:
            lib().dqmc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- configuration -----------------------------------------------------------------------
    def set_chain_mu(self, chain, mu):
        self._chk(lib().dqmc_set_chain_mu(self._h, chain, mu), "set_chain_mu")

    def set_collect(self, eqtime, dynamic):
        self._chk(lib().dqmc_set_collect(self._h, int(eqtime), int(dynamic)), "set_collect")

    def set_seed(self, chain, seed):
        self._chk(lib().dqmc_set_seed(self._h, chain, int(seed)), "set_seed")

    def rng_state(self, chain):
        x = C.c_uint64()
        self._chk(lib().dqmc_get_rng_state(self._h, chain, C.byref(x)), "get_rng_state")
        return int(x.value)

    def fields_random(self, chain):
        self._chk(lib().dqmc_fields_random(self._h, chain), "fields_random")

    def set_fields(self, chain, s):
        s = np.ascontiguousarray(np.asarray(s).reshape(self.nt, self.n), dtype=np.int8)
        self._chk(lib().dqmc_set_fields(self._h, chain, s.ctypes.data_as(C.POINTER(C.c_int8))), "set_fields")

    def fields(self, chain):
        s = np.zeros((self.nt, self.n), dtype=np.int8)
        self._chk(lib().dqmc_get_fields(self._h, chain, s.ctypes.data_as(C.POINTER(C.c_int8))), "get_fields")
        return s

    # --- observables on the device ---------------------------------------------------------------
    def measure_setup(self, obs_mask, nbins, kxy=None):
        """Select the DQMC_OBS_* observables, the bin count and the momentum list [(kx, ky), ...]."""
        k = np.ascontiguousarray(np.zeros((0, 2)) if kxy is None else kxy, dtype=np.float64).reshape(-1, 2)
        self._nk, self._nbins = k.shape[0], nbins
        self._chk(lib().dqmc_measure_setup(self._h, obs_mask, nbins, k.shape[0],
                                           k.ctypes.data_as(C.POINTER(C.c_double)) if k.size else None),
                  "measure_setup")

    def measure_equaltime(self):
        self._chk(lib().dqmc_measure_equaltime(self._h), "measure_equaltime")

    def measure_dynamic(self):
        self._chk(lib().dqmc_measure_dynamic(self._h), "measure_dynamic")

    def push_bin(self, b):
        self._chk(lib().dqmc_measure_push_bin(self._h, b), "measure_push_bin")

    def bins(self, chain, obs):
        """One chain's bins of one observable: [nbins], [nbins, nk, nt] or [nbins, nt]."""
        shape = {OBS_GREEN_KT: (self._nbins, self._nk, self.nt), OBS_LDOS: (self._nbins, self.nt),
                 OBS_SPIN: (self._nbins, self.nt)}.get(obs, (self._nbins,))
        out = np.zeros(shape)
        self._chk(lib().dqmc_get_bins(self._h, chain, obs, out.ctypes.data_as(C.POINTER(C.c_double))), "get_bins")
        return out

    def set_watchdog(self, max_dG):
        """Sweeps return DQMC_ENUMERIC when max|G_recomputed - G_working| exceeds max_dG (0 = off)."""
        self._chk(lib().dqmc_set_watchdog(self._h, float(max_dG)), "set_watchdog")

    def health(self, chain):
        """HEALTH_* flags one chain raised during the last call."""
        x = C.c_uint32()
        self._chk(lib().dqmc_get_health(self._h, chain, C.byref(x)), "get_health")
        return int(x.value)

    def use_graphs(self, on=True):
        self._chk(lib().dqmc_use_graphs(self._h, int(on)), "use_graphs")

    # --- sweep path ----------------------------------------------------------------------------
    def init(self):
        self._chk(lib().dqmc_init(self._h), "init")

    def set_thermalization(self, on):
        self._chk(lib().dqmc_set_thermalization(self._h, int(on)), "set_thermalization")

    def sweep_up(self):
        self._chk(lib().dqmc_sweep_up(self._h), "sweep_up")

    def sweep_down(self):
        self._chk(lib().dqmc_sweep_down(self._h), "sweep_down")

    def sweep_dynamic(self):
        self._chk(lib().dqmc_sweep_dynamic(self._h), "sweep_dynamic")

    def acceptance_ratio(self, chain, t, i):
        r = C.c_double()
        self._chk(lib().dqmc_acceptance_ratio(self._h, chain, t, i, C.byref(r)), "acceptance_ratio")
        return r.value

    def update_green_function(self, chain, t, i):
        self._chk(lib().dqmc_update_green_function(self._h, chain, t, i), "update_green_function")

    def flip_field(self, chain, t, i):
        self._chk(lib().dqmc_flip_field(self._h, chain, t, i), "flip_field")

    def apply_bop(self, chain, op, t, g):
        g = np.array(g, dtype=np.float64, order="F", copy=True)
        self._chk(lib().dqmc_apply_bop(self._h, chain, op, t, g.ctypes.data_as(C.POINTER(C.c_double))), "apply_bop")
        return g

    # --- results -------------------------------------------------------------------------------
    def gf(self, chain, kind=GF_CURRENT, t=0):
        out = np.zeros((self.n, self.n), order="F")
        self._chk(lib().dqmc_get_gf(self._h, chain, kind, t, out.ctypes.data_as(C.POINTER(C.c_double))), "get_gf")
        return out

    def set_gf(self, chain, g):
        g = np.asfortranarray(g, dtype=np.float64)
        self._chk(lib().dqmc_set_gf(self._h, chain, g.ctypes.data_as(C.POINTER(C.c_double))), "set_gf")

    def stats(self, chain):
        we, wd, dg = C.c_double(), C.c_double(), C.c_double()
        a, v = C.c_longlong(), C.c_longlong()
        self._chk(lib().dqmc_get_stats(self._h, chain, C.byref(we), C.byref(wd), C.byref(dg), C.byref(a),
                                       C.byref(v)), "get_stats")
        return {"wrap_eq": we.value, "wrap_dyn": wd.value, "max_abs_dG": dg.value, "accepts": a.value,
                "visits": v.value}

    def accept_mask(self, chain):
        out = np.zeros((self.nt, (self.n + 31) // 32), dtype=np.uint32)
        self._chk(lib().dqmc_get_accept_mask(self._h, chain, out.ctypes.data_as(C.POINTER(C.c_uint32))),
                  "get_accept_mask")
        return out

    def synchronize(self):
        self._chk(lib().dqmc_synchronize(self._h), "synchronize")

    def event_record(self, slot):
        self._chk(lib().dqmc_event_record(self._h, slot), "event_record")

    def kernel_timer_begin(self, kernel, max_launches):
        self._chk(lib().dqmc_kernel_timer_begin(self._h, kernel, max_launches), "kernel_timer_begin")

    def kernel_timer_end(self):
        ms = C.c_double()
        n = C.c_int()
        self._chk(lib().dqmc_kernel_timer_end(self._h, C.byref(ms), C.byref(n)), "kernel_timer_end")
        return float(ms.value), int(n.value)

    def tuning(self):
        """Autotuned GEMM tiles {'fp64': name, 'fp32': name, 'ms': {(prec, name): ms}}."""
        a, b = C.c_int(), C.c_int()
        ms = (C.c_float * (2 * len(TILE_NAMES)))()
        self._chk(lib().dqmc_get_tuning(self._h, C.byref(a), C.byref(b), ms), "get_tuning")
        nt = len(TILE_NAMES)
        return {"fp64": TILE_NAMES[a.value], "fp32": TILE_NAMES[b.value],
                "ms": {(p, TILE_NAMES[t]): ms[i * nt + t] for i, p in enumerate(("fp64", "fp32")) for t in range(nt)}}

    def forward_fixed(self):
        """(logdet [chains], gdiag [chains, nt, ns]) at the current fields, no updates."""
        ld = np.zeros(self.n_chains)
        gd = np.zeros((self.n_chains, self.nt, self.n))
        dp = C.POINTER(C.c_double)
        self._chk(lib().dqmc_forward_fixed(self._h, ld.ctypes.data_as(dp), gd.ctypes.data_as(dp)), "forward_fixed")
        return ld, gd

    def adjoint(self, gbar_logdet=None, gbar_gdiag=None):
        """[chains, 3] = cotangents contracted with d/d(nnt, mu, u) of forward_fixed's outputs."""
        dp = C.POINTER(C.c_double)
        gl = gg = None
        if gbar_logdet is not None:
            gl = np.ascontiguousarray(gbar_logdet, dtype=np.float64).reshape(self.n_chains)
        if gbar_gdiag is not None:
            gg = np.ascontiguousarray(gbar_gdiag, dtype=np.float64).reshape(self.n_chains, self.nt, self.n)
        out = np.zeros((self.n_chains, 3))
        self._chk(lib().dqmc_adjoint(self._h, gl.ctypes.data_as(dp) if gl is not None else None,
                                     gg.ctypes.data_as(dp) if gg is not None else None, out.ctypes.data_as(dp)),
                  "adjoint")
        return out

    def set_tuning(self, fp64, fp32):
        self._chk(lib().dqmc_set_tuning(self._h, TILE_NAMES.index(fp64), TILE_NAMES.index(fp32)), "set_tuning")

    def event_elapsed(self, a, b):
        ms = C.c_float()
        self._chk(lib().dqmc_event_elapsed(self._h, a, b, C.byref(ms)), "event_elapsed")
        return float(ms.value)
