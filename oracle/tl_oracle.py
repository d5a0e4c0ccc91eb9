"""ctypes view of the CPU oracle (oracle/_build/libtl_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or CPU baseline.  The product
package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TLO_ORACLE_LIB names another build of these sources (the tests that plant a
# mistake in the oracle and check that the pins catch it load such a build)
_LIB_PATH = os.environ.get("TLO_ORACLE_LIB") or os.path.join(_HERE, "_build", "libtl_oracle.so")
_lib = None

P = C.c_void_p
I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
I8P = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
U8P = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
U32P = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


class _Model(C.Structure):
    _fields_ = [
        ("N", C.c_int32), ("M", C.c_int32), ("nnz", C.c_int32),
        ("col_ptr", C.POINTER(C.c_int32)), ("row_ind", C.POINTER(C.c_int32)),
        ("stoich", C.POINTER(C.c_int8)), ("rates", C.POINTER(C.c_double)),
        ("rates_f", C.POINTER(C.c_float)), ("x0", C.POINTER(C.c_int32)),
        ("order", C.POINTER(C.c_int8)), ("s0", C.POINTER(C.c_int32)),
        ("s1", C.POINTER(C.c_int32)), ("row_ptr", C.POINTER(C.c_int32)),
        ("csr_packed", C.POINTER(C.c_int32)), ("g_order", C.POINTER(C.c_int8)),
        ("s2", C.POINTER(C.c_int32)),
    ]


def _poisson(p):
    """Sampler flags: 'normal' (a Poisson count by inversion below mean 80,
    the reference's normal approximation from 80), 'exact' (PTRS from 80),
    'fast' (PTRS from 10 to 80), 'fast-exact' (PTRS from 10)."""
    v = {"normal": 0, "exact": 1, "ptrs": 1, "fast": 2, "fast-exact": 3}.get(p, p)
    if v not in (0, 1, 2, 3):
        raise ValueError(f"unknown Poisson sampler {p!r}")
    return v


class _Params(C.Structure):
    _fields_ = [
        ("t_final", C.c_double), ("eps", C.c_double), ("seed", C.c_uint64),
        ("model_id", C.c_uint32), ("precision", C.c_int32), ("max_iter", C.c_int64),
        ("sensitivities", C.c_int32), ("n_intervals", C.c_int32), ("n_track", C.c_int32),
        ("track", C.POINTER(C.c_int32)), ("n_threads", C.c_int32), ("poisson", C.c_int32),
    ]


class _Outputs(C.Structure):
    _fields_ = [
        ("x_final", C.POINTER(C.c_int32)), ("score", C.POINTER(C.c_double)),
        ("grid", C.POINTER(C.c_double)), ("steps", C.POINTER(C.c_int64)),
        ("ssa_events", C.POINTER(C.c_int64)), ("rejections", C.POINTER(C.c_int64)),
        ("status", C.POINTER(C.c_int32)),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.tlo_philox4x32_10.argtypes = [U32P, U32P, U32P]
        L.tlo_model_create.argtypes = [C.c_int32, C.c_int32, I32P, I32P, I8P, F64P, I32P,
                                       C.POINTER(C.POINTER(_Model))]
        L.tlo_model_create.restype = C.c_int
        L.tlo_model_destroy.argtypes = [C.POINTER(_Model)]
        L.tlo_model_set_net_stoichiometry.argtypes = [C.POINTER(_Model)]
        L.tlo_model_set_net_stoichiometry.restype = C.c_int
        for suf, RP in (("f64", F64P), ("f32", F32P)):
            getattr(L, "tlo_propensities_" + suf).argtypes = [C.POINTER(_Model), C.c_int64, I32P, RP, U8P, F64P, F64P]
            getattr(L, "tlo_tau_" + suf).argtypes = [C.POINTER(_Model), C.c_int64, C.c_double, I32P, RP, U8P, F64P, I32P, I32P]
            getattr(L, "tlo_select_" + suf).argtypes = [C.POINTER(_Model), C.c_int64, C.c_uint64, C.c_uint32, C.c_int64,
                                                       U32P, RP, U8P, F64P, F64P, F64P, F64P, I32P, U8P]
            getattr(L, "tlo_fire_" + suf).argtypes = [C.POINTER(_Model), C.c_int64, C.c_uint64, C.c_uint32, C.c_int64,
                                                     U32P, RP, U8P, F64P, I32P, I32P]
            getattr(L, "tlo_ssa_" + suf).argtypes = [C.POINTER(_Model), C.c_int64, C.c_uint64, C.c_uint32, C.c_int64,
                                                    I32P, F64P, U32P, C.c_double, C.c_int32, I32P]
        L.tlo_poisson_f64.argtypes = [C.c_double, C.c_uint64, C.c_uint32, C.c_int64, C.c_uint32, C.c_int32, C.c_int32]
        L.tlo_poisson_f64.restype = C.c_int32
        L.tlo_poisson_f32.argtypes = [C.c_float, C.c_uint64, C.c_uint32, C.c_int64, C.c_uint32, C.c_int32, C.c_int32]
        L.tlo_poisson_f32.restype = C.c_int32
        L.tlo_poisson_ptrs.argtypes = [C.c_double, C.c_uint64, C.c_uint32, C.c_int64, C.c_uint32, C.c_int32]
        L.tlo_poisson_ptrs.restype = C.c_int32
        L.tlo_poisson_ptrs_many.argtypes = [C.c_double, C.c_uint64, C.c_uint32, C.c_int64, C.c_int64, C.c_uint32,
                                            C.c_int32, I32P]
        L.tlo_update.argtypes = [C.POINTER(_Model), C.c_int64, I32P, I32P, I32P, U8P]
        L.tlo_simulate.argtypes = [C.POINTER(_Model), C.POINTER(_Params), C.c_int64, C.c_int64, C.POINTER(_Outputs)]
        L.tlo_simulate.restype = C.c_int
        _lib = L
    return _lib


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().tlo_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def _rtype(precision):
    return np.float64 if precision in (0, "fp64", "f64") else np.float32


def _suf(precision):
    return "f64" if _rtype(precision) == np.float64 else "f32"


class OracleModel:
    """Model built from channel-major (CSC) entries, as the product's
    tl_model_from_csc takes them."""

    def __init__(self, N, M, col_ptr, row_ind, stoich, rates, x0, net_stoichiometry=False):
        self.N, self.M = int(N), int(M)
        ptr = C.POINTER(_Model)()
        st = lib().tlo_model_create(self.N, self.M, np.ascontiguousarray(col_ptr, np.int32),
                                    np.ascontiguousarray(row_ind, np.int32),
                                    np.ascontiguousarray(stoich, np.int8),
                                    np.ascontiguousarray(rates, np.float64),
                                    np.ascontiguousarray(x0, np.int32), C.byref(ptr))
        if st != 0:
            raise ValueError(f"tlo_model_create failed with status {st}")
        self._p = ptr
        if net_stoichiometry:
            st = lib().tlo_model_set_net_stoichiometry(ptr)
            if st != 0:
                raise ValueError(f"tlo_model_set_net_stoichiometry failed with status {st}")
        m = ptr.contents
        self.nnz = m.nnz
        self.order = np.ctypeslib.as_array(m.order, (self.M,)).copy()
        self.s0 = np.ctypeslib.as_array(m.s0, (self.M,)).copy()
        self.s1 = np.ctypeslib.as_array(m.s1, (self.M,)).copy()
        self.row_ptr = np.ctypeslib.as_array(m.row_ptr, (self.N + 1,)).copy()
        nz = int(self.row_ptr[-1])
        self.csr_packed = np.ctypeslib.as_array(m.csr_packed, (max(self.nnz, 1),))[:nz].copy()
        self.x0 = np.ascontiguousarray(x0, np.int32).copy()

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.tlo_model_destroy(self._p)
            self._p = None

    # ---- stages on SoA buffers ----
    def propensities(self, x, precision="fp64"):
        x = np.ascontiguousarray(x, np.int32)
        R = x.shape[1]
        a = np.zeros((self.M, R), _rtype(precision))
        crit = np.zeros((self.M, R), np.uint8)
        a0 = np.zeros(R)
        a0c = np.zeros(R)
        getattr(lib(), "tlo_propensities_" + _suf(precision))(self._p, R, x, a, crit, a0, a0c)
        return a, crit, a0, a0c

    def tau(self, x, a, crit, eps, precision="fp64"):
        R = x.shape[1]
        tau1 = np.zeros(R)
        am = np.zeros(R, np.int32)
        wh = np.zeros(R, np.int32)
        getattr(lib(), "tlo_tau_" + _suf(precision))(self._p, R, float(eps), np.ascontiguousarray(x, np.int32),
                                                      np.ascontiguousarray(a, _rtype(precision)),
                                                      np.ascontiguousarray(crit, np.uint8), tau1, am, wh)
        return tau1, am, wh

    def select(self, seed, epoch, a, crit, a0c, tau1, t_rem, precision="fp64", model_id=0, rbase=0):
        R = a.shape[1]
        tau = np.zeros(R)
        jc = np.zeros(R, np.int32)
        cl = np.zeros(R, np.uint8)
        getattr(lib(), "tlo_select_" + _suf(precision))(
            self._p, R, int(seed), int(model_id), int(rbase), np.ascontiguousarray(epoch, np.uint32),
            np.ascontiguousarray(a, _rtype(precision)), np.ascontiguousarray(crit, np.uint8),
            np.ascontiguousarray(a0c, np.float64), np.ascontiguousarray(tau1, np.float64),
            np.ascontiguousarray(t_rem, np.float64), tau, jc, cl)
        return tau, jc, cl

    def fire(self, seed, epoch, a, crit, tau, jc, precision="fp64", model_id=0, rbase=0):
        R = a.shape[1]
        k = np.zeros((self.M, R), np.int32)
        getattr(lib(), "tlo_fire_" + _suf(precision))(
            self._p, R, int(seed), int(model_id), int(rbase), np.ascontiguousarray(epoch, np.uint32),
            np.ascontiguousarray(a, _rtype(precision)), np.ascontiguousarray(crit, np.uint8),
            np.ascontiguousarray(tau, np.float64), np.ascontiguousarray(jc, np.int32), k)
        return k

    def update(self, x, k):
        R = x.shape[1]
        xo = np.zeros((self.N, R), np.int32)
        neg = np.zeros(R, np.uint8)
        lib().tlo_update(self._p, R, np.ascontiguousarray(x, np.int32), np.ascontiguousarray(k, np.int32), xo, neg)
        return xo, neg

    def ssa(self, seed, x, t, epoch, t_final, n_events, precision="fp64", model_id=0, rbase=0):
        x = np.array(x, np.int32, copy=True, order="C")
        t = np.array(t, np.float64, copy=True)
        epoch = np.array(epoch, np.uint32, copy=True)
        R = x.shape[1]
        ev = np.zeros(R, np.int32)
        getattr(lib(), "tlo_ssa_" + _suf(precision))(self._p, R, int(seed), int(model_id), int(rbase), x, t, epoch,
                                                      float(t_final), int(n_events), ev)
        return x, t, epoch, ev

    # ---- full simulation ----
    def simulate(self, R, t_final, eps=0.03, seed=1234, precision="fp64", max_iter=0, sensitivities=False,
                 n_intervals=0, track=None, n_threads=0, model_id=0, rbase=0, poisson="normal"):
        R = int(R)
        track = np.arange(self.N, dtype=np.int32) if track is None else np.ascontiguousarray(track, np.int32)
        p = _Params(float(t_final), float(eps), int(seed), int(model_id), 0 if _rtype(precision) == np.float64 else 1,
                    int(max_iter), int(bool(sensitivities)), int(n_intervals), int(track.size),
                    track.ctypes.data_as(C.POINTER(C.c_int32)), int(n_threads), _poisson(poisson))
        xf = np.zeros((R, self.N), np.int32)
        score = np.zeros((R, self.M)) if sensitivities else None
        grid = np.zeros((R, n_intervals + 1, track.size)) if n_intervals > 0 else None
        steps = np.zeros(R, np.int64)
        ssa = np.zeros(R, np.int64)
        rej = np.zeros(R, np.int64)
        status = np.zeros(R, np.int32)

        def ptr(arr, ct):
            return arr.ctypes.data_as(C.POINTER(ct)) if arr is not None else C.POINTER(ct)()

        o = _Outputs(ptr(xf, C.c_int32), ptr(score, C.c_double), ptr(grid, C.c_double), ptr(steps, C.c_int64),
                     ptr(ssa, C.c_int64), ptr(rej, C.c_int64), ptr(status, C.c_int32))
        nt = lib().tlo_simulate(self._p, C.byref(p), int(rbase), R, C.byref(o))
        return dict(x_final=xf, score=score, grid=grid, steps=steps, ssa_events=ssa, rejections=rej,
                    status=status, threads=nt)

    def poisson_ptrs(self, mean, seed, rid, epoch, channel, model_id=0, n=None):
        """PTRS draw of channel `channel` at `epoch` for realization rid, or an
        array over realizations rid .. rid + n - 1."""
        if n is None:
            return lib().tlo_poisson_ptrs(float(mean), int(seed), int(model_id), int(rid), int(epoch), int(channel))
        out = np.zeros(int(n), np.int32)
        lib().tlo_poisson_ptrs_many(float(mean), int(seed), int(model_id), int(rid), int(n), int(epoch), int(channel),
                                    out)
        return out

    def poisson(self, mean, seed, rid, epoch, channel, precision="fp64", model_id=0, poisson="normal"):
        f = lib().tlo_poisson_f64 if _rtype(precision) == np.float64 else lib().tlo_poisson_f32
        return f(mean, int(seed), int(model_id), int(rid), int(epoch), int(channel), _poisson(poisson))
