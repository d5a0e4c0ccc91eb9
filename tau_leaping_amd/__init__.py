"""MI355X tau-leaping Gillespie engine.

Python view of the C ABI in ``include/tl_api.h`` (libtau_leaping_amd.so, HIP
for gfx950).  The library has no CPU path: simulation calls need an AMD GPU
and raise if the library or the device is missing.

The surface mirrors the reference's simulation path
(tau_leaping/model.h:41-55, tau_leaping/tau_leaping.cpp:40-190):
``Model.from_stochkit`` is ``Model::ReadXML``; ``Model.simulate`` runs the
realizations that the reference runs one by one through ``PerformCalc`` and
returns the statistics ``statistics.cpp`` would write.  ``mean_final_state``
wraps the likelihood-ratio sensitivities as a ``torch.autograd.Function``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import models  # noqa: F401  (reaction networks and generators)

__all__ = ["Model", "SimResult", "simulate", "batch_simulate", "mean_final_state", "lib", "TLError",
           "stage_propensities", "stage_tau", "stage_select", "stage_fire", "stage_update", "stage_ssa",
           "FP64", "FP32", "MODE_AUTO", "MODE_PLAIN", "MODE_PERSISTENT", "POISSON_NORMAL", "POISSON_EXACT",
           "POISSON_FAST", "POISSON_FAST_EXACT", "models"]

FP64, FP32 = 0, 1
MODE_AUTO, MODE_PLAIN, MODE_PERSISTENT = 0, 1, 2
POISSON_NORMAL, POISSON_EXACT, POISSON_FAST, POISSON_FAST_EXACT = 0, 1, 2, 3
BLOCK_TUNE = -1  # block_size: a timed choice among the launch shapes, cached per size class

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtau_leaping_amd.so")
_lib = None


class TLError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"tau_leaping_amd error {status}: {msg}")
        self.status = status


class _Params(C.Structure):
    _fields_ = [("t_final", C.c_double), ("eps", C.c_double), ("n_realizations", C.c_int64),
                ("seed", C.c_uint64), ("max_iter", C.c_int64), ("precision", C.c_int32),
                ("n_track", C.c_int32), ("track", C.POINTER(C.c_int32)), ("n_intervals", C.c_int32),
                ("sensitivities", C.c_int32), ("device", C.c_int32), ("mode", C.c_int32),
                ("block_size", C.c_int32), ("model_id", C.c_uint32), ("realization_offset", C.c_int64),
                ("poisson", C.c_int32)]


class _Result(C.Structure):
    _fields_ = [("mean", C.POINTER(C.c_double)), ("stddev", C.POINTER(C.c_double)),
                ("grid_times", C.POINTER(C.c_double)), ("final_x", C.POINTER(C.c_int32)),
                ("dmean_dc", C.POINTER(C.c_double)), ("dmean_dc_se", C.POINTER(C.c_double)),
                ("steps", C.POINTER(C.c_int64)), ("score", C.POINTER(C.c_double)),
                ("total_steps", C.c_int64), ("total_ssa_events", C.c_int64), ("total_rejections", C.c_int64),
                ("n_failed", C.c_int64), ("first_failure", C.c_int32), ("block_size", C.c_int32),
                ("sim_seconds", C.c_double), ("upload_seconds", C.c_double), ("kernel_ms", C.c_double),
                ("rejections", C.POINTER(C.c_int64)), ("status", C.POINTER(C.c_int32)),
                ("grid", C.POINTER(C.c_double)), ("form", C.c_int32)]


_SIGS = {
    "tl_model_from_stochkit": ([C.c_char_p, C.POINTER(C.c_void_p)], C.c_int32),
    "tl_model_from_csc": ([C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                           C.POINTER(C.c_void_p)], C.c_int32),
    "tl_model_destroy": ([C.c_void_p], None),
    "tl_model_n_species": ([C.c_void_p], C.c_int32),
    "tl_model_n_channels": ([C.c_void_p], C.c_int32),
    "tl_model_nnz": ([C.c_void_p], C.c_int32),
    "tl_model_species_name": ([C.c_void_p, C.c_int32], C.c_char_p),
    "tl_model_species_index": ([C.c_void_p, C.c_char_p, C.POINTER(C.c_int32)], C.c_int32),
    "tl_model_get_rates": ([C.c_void_p, C.c_void_p], C.c_int32),
    "tl_model_set_rates": ([C.c_void_p, C.c_void_p], C.c_int32),
    "tl_model_get_x0": ([C.c_void_p, C.c_void_p], C.c_int32),
    "tl_model_set_x0": ([C.c_void_p, C.c_void_p], C.c_int32),
    "tl_model_structure": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int32),
    "tl_model_set_net_stoichiometry": ([C.c_void_p, C.c_int32], C.c_int32),
    "tl_model_net_stoichiometry": ([C.c_void_p], C.c_int32),
    "tl_model_layout_nnz": ([C.c_void_p], C.c_int32),
    "tl_sim_params_default": ([C.POINTER(_Params)], None),
    "tl_simulate": ([C.c_void_p, C.POINTER(_Params), C.POINTER(_Result)], C.c_int32),
    "tl_batch_simulate": ([C.POINTER(C.c_void_p), C.c_int32, C.POINTER(_Params), C.POINTER(_Result), C.c_int32],
                          C.c_int32),
    "tl_stage_propensities": ([C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p], C.c_int32),
    "tl_stage_tau": ([C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                      C.c_void_p, C.c_void_p, C.c_void_p], C.c_int32),
    "tl_stage_select": ([C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_uint32, C.c_int64, C.c_void_p,
                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                         C.c_void_p], C.c_int32),
    "tl_stage_fire": ([C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_uint32, C.c_int64, C.c_void_p,
                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int32),
    "tl_stage_update": ([C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int32),
    "tl_stage_ssa": ([C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_uint32, C.c_int64, C.c_void_p,
                      C.c_void_p, C.c_void_p, C.c_double, C.c_int32, C.c_void_p], C.c_int32),
    "tl_simulate_noalloc": ([C.c_void_p, C.POINTER(_Params), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
                            C.c_int32),
    "tl_last_error": ([], C.c_char_p),
    "tl_version": ([], C.c_char_p),
    "tl_device_count": ([C.POINTER(C.c_int32)], C.c_int32),
}


def lib():
    """Load libtau_leaping_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {os.path.join(_HERE, 'csrc')}` "
                               "or __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(st):
    if st != 0:
        raise TLError(st, lib().tl_last_error().decode(errors="replace"))


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _prec(p):
    if p in (FP64, "fp64", "f64", "double", np.float64):
        return FP64
    if p in (FP32, "fp32", "f32", "float", np.float32):
        return FP32
    raise ValueError(f"unknown precision {p!r}")


def _poisson(p):
    """Firing-count sampler: 'normal' (default: a Poisson count by inversion
    below mean 80, the reference's |N(m, m)| from 80), 'exact' (PTRS from 80),
    'fast' (PTRS from 10 to 80), 'fast-exact' (PTRS from 10)."""
    v = {"normal": POISSON_NORMAL, "exact": POISSON_EXACT, "fast": POISSON_FAST,
         "fast-exact": POISSON_FAST_EXACT}.get(p, p)
    if v not in (POISSON_NORMAL, POISSON_EXACT, POISSON_FAST, POISSON_FAST_EXACT):
        raise ValueError(f"unknown Poisson sampler {p!r}")
    return v


def _mode(m):
    return {"auto": MODE_AUTO, "plain": MODE_PLAIN, "persistent": MODE_PERSISTENT}.get(m, m)


@dataclass
class SimResult:
    mean: np.ndarray           # [n_grid][n_track]
    stddev: np.ndarray         # [n_grid][n_track]
    grid_times: np.ndarray     # [n_grid]
    track: np.ndarray          # species indices
    final_x: Optional[np.ndarray]   # [R][n_track]
    dmean_dc: Optional[np.ndarray]  # [n_track][M]
    dmean_dc_se: Optional[np.ndarray]
    steps: np.ndarray          # [R]
    score: Optional[np.ndarray]     # [R][M] per-realization likelihood-ratio score (return_scores=True)
    total_steps: int
    total_ssa_events: int
    total_rejections: int
    n_failed: int
    first_failure: int
    block_size: int
    sim_seconds: float
    upload_seconds: float
    kernel_ms: float
    rejections: np.ndarray     # [R] rejected leap attempts per realization
    status: np.ndarray         # [R] final status per realization
    grid: Optional[np.ndarray]  # [R][n_grid][n_track] per-realization time grid (n_intervals > 0)
    form: str                  # "plain" or "persistent": the form that ran

    @property
    def leaps_per_second(self) -> float:
        """Accepted realization-steps per second of the simulation kernel(s)."""
        return self.total_steps / (self.kernel_ms * 1e-3) if self.kernel_ms > 0 else float("nan")


class Model:
    """A reaction network on the engine (library-owned handle)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value and _lib is not None:
            _lib.tl_model_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @classmethod
    def from_stochkit(cls, path: str) -> "Model":
        h = C.c_void_p()
        _check(lib().tl_model_from_stochkit(os.fsencode(path), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_csc(cls, n_species, col_ptr, row_ind, stoich, rates, x0) -> "Model":
        col_ptr = np.ascontiguousarray(col_ptr, np.int32)
        row_ind = np.ascontiguousarray(row_ind, np.int32)
        stoich = np.ascontiguousarray(stoich, np.int8)
        rates = np.ascontiguousarray(rates, np.float64)
        x0 = np.ascontiguousarray(x0, np.int32)
        h = C.c_void_p()
        _check(lib().tl_model_from_csc(int(n_species), int(col_ptr.size - 1), _ptr(col_ptr), _ptr(row_ind),
                                       _ptr(stoich), _ptr(rates), _ptr(x0), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_network(cls, net) -> "Model":
        return cls.from_csc(net.n_species, *net.to_csc())

    @property
    def n_species(self) -> int:
        return lib().tl_model_n_species(self._h)

    @property
    def n_channels(self) -> int:
        return lib().tl_model_n_channels(self._h)

    @property
    def nnz(self) -> int:
        return lib().tl_model_nnz(self._h)

    @property
    def species_names(self) -> List[str]:
        return [lib().tl_model_species_name(self._h, i).decode() for i in range(self.n_species)]

    def species_index(self, name: str) -> int:
        out = C.c_int32()
        _check(lib().tl_model_species_index(self._h, name.encode(), C.byref(out)))
        return out.value

    @property
    def rates(self) -> np.ndarray:
        r = np.zeros(self.n_channels)
        _check(lib().tl_model_get_rates(self._h, _ptr(r)))
        return r

    @rates.setter
    def rates(self, v):
        v = np.ascontiguousarray(v, np.float64)
        if v.size != self.n_channels:
            raise ValueError("rates must have one entry per channel")
        _check(lib().tl_model_set_rates(self._h, _ptr(v)))

    @property
    def x0(self) -> np.ndarray:
        x = np.zeros(self.n_species, np.int32)
        _check(lib().tl_model_get_x0(self._h, _ptr(x)))
        return x

    @x0.setter
    def x0(self, v):
        v = np.ascontiguousarray(v, np.int32)
        if v.size != self.n_species:
            raise ValueError("x0 must have one entry per species")
        _check(lib().tl_model_set_x0(self._h, _ptr(v)))

    @property
    def net_stoichiometry(self) -> bool:
        """Whether the tau bound takes a catalyst's entries as one net entry
        (off: each entry as read, as the reference builds them)."""
        return bool(lib().tl_model_net_stoichiometry(self._h))

    @net_stoichiometry.setter
    def net_stoichiometry(self, on):
        _check(lib().tl_model_set_net_stoichiometry(self._h, 1 if on else 0))

    @property
    def layout_nnz(self) -> int:
        """Entries of the packed layouts (fewer than nnz with net stoichiometry)."""
        return lib().tl_model_layout_nnz(self._h)

    def structure(self):
        """(order, s0, s1, row_ptr, csr_packed) as the engine derived them."""
        M, N, Z = self.n_channels, self.n_species, self.layout_nnz
        order = np.zeros(M, np.int8)
        s0 = np.zeros(M, np.int32)
        s1 = np.zeros(M, np.int32)
        rp = np.zeros(N + 1, np.int32)
        csr = np.zeros(max(Z, 1), np.int32)
        _check(lib().tl_model_structure(self._h, _ptr(order), _ptr(s0), _ptr(s1), _ptr(rp), _ptr(csr)))
        return order, s0, s1, rp, csr[:Z]

    def params(self, t_final, n_realizations, eps=0.03, seed=1234, precision="fp64", track=None, n_intervals=0,
               sensitivities=False, max_iter=0, mode="auto", block_size=0, device=0, model_id=0,
               realization_offset=0, poisson="normal"):
        p = _Params()
        lib().tl_sim_params_default(C.byref(p))
        p.t_final = float(t_final)
        p.eps = float(eps)
        p.n_realizations = int(n_realizations)
        p.seed = int(seed)
        p.max_iter = int(max_iter)
        p.precision = _prec(precision)
        trk = None if track is None else np.ascontiguousarray(track, np.int32)
        if trk is not None:
            p.n_track = int(trk.size)
            p.track = trk.ctypes.data_as(C.POINTER(C.c_int32))
        p.n_intervals = int(n_intervals)
        p.sensitivities = int(bool(sensitivities))
        p.device = int(device)
        p.mode = _mode(mode)
        p.block_size = int(block_size)
        p.model_id = int(model_id)
        p.realization_offset = int(realization_offset)
        p.poisson = _poisson(poisson)
        return p, trk

    def simulate(self, t_final, n_realizations, **kw) -> SimResult:
        return batch_simulate([self], [dict(t_final=t_final, n_realizations=n_realizations, **kw)],
                              mode=kw.get("mode", "auto"))[0]


def _result_buffers(model: Model, p: _Params, trk, want_final=True, want_score=False, want_grid=False):
    T = int(p.n_track) if p.n_track > 0 else model.n_species
    G = p.n_intervals + 1 if p.n_intervals > 0 else 1
    R = int(p.n_realizations)
    bufs = dict(mean=np.zeros((G, T)), stddev=np.zeros((G, T)), grid_times=np.zeros(G),
                final_x=np.zeros((R, T), np.int32) if want_final else None,
                dmean_dc=np.zeros((T, model.n_channels)) if p.sensitivities else None,
                dmean_dc_se=np.zeros((T, model.n_channels)) if p.sensitivities else None,
                steps=np.zeros(R, np.int64),
                score=np.zeros((R, model.n_channels)) if (want_score and p.sensitivities) else None,
                rejections=np.zeros(R, np.int64), status=np.zeros(R, np.int32),
                grid=np.zeros((R, G, T)) if (want_grid and p.n_intervals > 0) else None)
    r = _Result()

    def P(a, ct):
        return a.ctypes.data_as(C.POINTER(ct)) if a is not None else C.POINTER(ct)()

    r.mean = P(bufs["mean"], C.c_double)
    r.stddev = P(bufs["stddev"], C.c_double)
    r.grid_times = P(bufs["grid_times"], C.c_double)
    r.final_x = P(bufs["final_x"], C.c_int32)
    r.dmean_dc = P(bufs["dmean_dc"], C.c_double)
    r.dmean_dc_se = P(bufs["dmean_dc_se"], C.c_double)
    r.steps = P(bufs["steps"], C.c_int64)
    r.score = P(bufs["score"], C.c_double)
    r.rejections = P(bufs["rejections"], C.c_int64)
    r.status = P(bufs["status"], C.c_int32)
    r.grid = P(bufs["grid"], C.c_double)
    track = trk if trk is not None else np.arange(model.n_species, dtype=np.int32)
    return r, bufs, track


_FORMS = {MODE_PLAIN: "plain", MODE_PERSISTENT: "persistent"}


def _to_result(r: _Result, bufs, track) -> SimResult:
    return SimResult(mean=bufs["mean"], stddev=bufs["stddev"], grid_times=bufs["grid_times"], track=track,
                     final_x=bufs["final_x"], dmean_dc=bufs["dmean_dc"], dmean_dc_se=bufs["dmean_dc_se"],
                     steps=bufs["steps"], score=bufs["score"], total_steps=r.total_steps, total_ssa_events=r.total_ssa_events,
                     total_rejections=r.total_rejections, n_failed=r.n_failed, first_failure=r.first_failure,
                     block_size=r.block_size, sim_seconds=r.sim_seconds, upload_seconds=r.upload_seconds,
                     kernel_ms=r.kernel_ms, rejections=bufs["rejections"], status=bufs["status"], grid=bufs["grid"],
                     form=_FORMS.get(r.form, "unknown"))


def batch_simulate(model_list: Sequence[Model], kwargs_list: Sequence[dict], mode="auto",
                   want_final=True) -> List[SimResult]:
    """Run many models; ``mode='persistent'`` is one work-stealing launch for
    the whole batch (a model too large for LDS keeps its realizations' state
    in HBM), ``'plain'`` the per-model stage-kernel loop, ``'auto'`` (default)
    persistent, plain for a model no persistent launch shape can run and for a
    model too large for LDS on which a timed trial of the first steps finds
    the plain form more than 15 % faster (neither when sensitivities or a time
    grid are asked for)."""
    n = len(model_list)
    params = (_Params * n)()
    keep, outs = [], []
    results = (_Result * n)()
    for i, (m, kw) in enumerate(zip(model_list, kwargs_list)):
        kw = dict(kw)
        kw.pop("mode", None)
        want_score = bool(kw.pop("return_scores", False))
        want_grid = bool(kw.pop("return_grid", False))
        p, trk = m.params(mode=mode, **kw)
        params[i] = p
        keep.append(trk)
        r, bufs, track = _result_buffers(m, p, trk, want_final, want_score, want_grid)
        results[i] = r
        outs.append((bufs, track))
    handles = (C.c_void_p * n)(*[m.handle.value for m in model_list])
    _check(lib().tl_batch_simulate(handles, n, params, results, _mode(mode)))
    return [_to_result(results[i], *outs[i]) for i in range(n)]


def simulate(model: Model, t_final, n_realizations, **kw) -> SimResult:
    return model.simulate(t_final, n_realizations, **kw)


# ------------------------------------------------------------------ stages --
def _dptr(t):
    return C.c_void_p(t.data_ptr())


def _need_cuda(*ts):
    for t in ts:
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError("stage buffers must be contiguous device tensors")


def stage_propensities(model: Model, x, precision="fp64"):
    import torch
    _need_cuda(x)
    N, R = x.shape
    dt = torch.float64 if _prec(precision) == FP64 else torch.float32
    a = torch.empty((model.n_channels, R), dtype=dt, device=x.device)
    crit = torch.empty((model.n_channels, R), dtype=torch.uint8, device=x.device)
    a0 = torch.empty(R, dtype=torch.float64, device=x.device)
    a0c = torch.empty(R, dtype=torch.float64, device=x.device)
    _check(lib().tl_stage_propensities(model.handle, _prec(precision), R, _dptr(x), _dptr(a), _dptr(crit),
                                       _dptr(a0), _dptr(a0c)))
    return a, crit, a0, a0c


def stage_tau(model: Model, x, a, crit, eps, precision="fp64"):
    import torch
    _need_cuda(x, a, crit)
    R = x.shape[1]
    tau1 = torch.empty(R, dtype=torch.float64, device=x.device)
    am = torch.empty(R, dtype=torch.int32, device=x.device)
    wh = torch.empty(R, dtype=torch.int32, device=x.device)
    _check(lib().tl_stage_tau(model.handle, _prec(precision), R, float(eps), _dptr(x), _dptr(a), _dptr(crit),
                              _dptr(tau1), _dptr(am), _dptr(wh)))
    return tau1, am, wh


def stage_select(model: Model, seed, epoch, a, crit, a0c, tau1, t_rem, precision="fp64", model_id=0, rbase=0):
    import torch
    _need_cuda(epoch, a, crit, a0c, tau1, t_rem)
    R = a.shape[1]
    tau = torch.empty(R, dtype=torch.float64, device=a.device)
    jc = torch.empty(R, dtype=torch.int32, device=a.device)
    cl = torch.empty(R, dtype=torch.uint8, device=a.device)
    _check(lib().tl_stage_select(model.handle, _prec(precision), R, int(seed), int(model_id), int(rbase),
                                 _dptr(epoch), _dptr(a), _dptr(crit), _dptr(a0c), _dptr(tau1), _dptr(t_rem),
                                 _dptr(tau), _dptr(jc), _dptr(cl)))
    return tau, jc, cl


def stage_fire(model: Model, seed, epoch, a, crit, tau, jc, precision="fp64", model_id=0, rbase=0):
    import torch
    _need_cuda(epoch, a, crit, tau, jc)
    R = a.shape[1]
    k = torch.empty((model.n_channels, R), dtype=torch.int32, device=a.device)
    _check(lib().tl_stage_fire(model.handle, _prec(precision), R, int(seed), int(model_id), int(rbase),
                               _dptr(epoch), _dptr(a), _dptr(crit), _dptr(tau), _dptr(jc), _dptr(k)))
    return k


def stage_update(model: Model, x, k):
    import torch
    _need_cuda(x, k)
    R = x.shape[1]
    xo = torch.empty_like(x)
    neg = torch.empty(R, dtype=torch.uint8, device=x.device)
    _check(lib().tl_stage_update(model.handle, R, _dptr(x), _dptr(k), _dptr(xo), _dptr(neg)))
    return xo, neg


def stage_ssa(model: Model, seed, x, t, epoch, t_final, n_events, precision="fp64", model_id=0, rbase=0):
    """In place on x, t, epoch; returns the events fired per realization."""
    import torch
    _need_cuda(x, t, epoch)
    R = x.shape[1]
    ev = torch.empty(R, dtype=torch.int32, device=x.device)
    _check(lib().tl_stage_ssa(model.handle, _prec(precision), R, int(seed), int(model_id), int(rbase), _dptr(x),
                              _dptr(t), _dptr(epoch), float(t_final), int(n_events), _dptr(ev)))
    return ev


# ------------------------------------------------------------- autograd ----
def _mean_final_fn():
    import torch

    class MeanFinalState(torch.autograd.Function):
        """E[X_track(t_final)] as a function of the rate constants.  Backward
        uses the likelihood-ratio (score-function) estimate dE[X_s]/dc_j the
        engine accumulates alongside the forward leaps."""

        @staticmethod
        def forward(ctx, rates, model, kw):
            # simulate at the given rates, then give the model its own back
            saved = model.rates
            model.rates = rates.detach().to("cpu", torch.float64).numpy()
            try:
                res = model.simulate(sensitivities=True, **kw)
            finally:
                model.rates = saved
            ctx.save_for_backward(torch.from_numpy(res.dmean_dc))
            ctx.rates_dtype = rates.dtype
            ctx.rates_device = rates.device
            return torch.from_numpy(res.mean[-1].copy()).to(rates.device, rates.dtype)

        @staticmethod
        def backward(ctx, g):
            (J,) = ctx.saved_tensors
            grad = g.detach().to("cpu", torch.float64) @ J
            return grad.to(ctx.rates_device, ctx.rates_dtype), None, None

    return MeanFinalState


_MFS = None


def mean_final_state(rates, model: Model, **kw):
    """Differentiable E[X_track(t_final)] (see MeanFinalState)."""
    global _MFS
    if _MFS is None:
        _MFS = _mean_final_fn()
    return _MFS.apply(rates, model, kw)
