"""tl_simulate_noalloc (the entry point that allocates no device memory)
through the C ABI against tl_simulate's persistent form on the same inputs:
the same kernel and launch shape, so final states, steps and statuses are
identical, bit for bit.  Two calls in a row, the second with fewer
realizations (the shared scratch keeps the first call's contents), for a
one-wavefront model and a block-group model, in FP64 and FP32; a model whose
constants are not on the device is refused."""
import ctypes as C

import numpy as np
import pytest

import tau_leaping_amd as T
from tau_leaping_amd import models

pytestmark = pytest.mark.gpu


def _noalloc(m, t_final, R, track, **kw):
    import torch

    p, trk = m.params(t_final, R, track=track, mode="persistent", **kw)
    fx = torch.full((R, len(track)), -1, dtype=torch.int32, device="cuda:0")
    steps = torch.full((R,), -1, dtype=torch.int64, device="cuda:0")
    status = torch.full((R,), -1, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    st = T.lib().tl_simulate_noalloc(m.handle, C.byref(p), C.c_void_p(fx.data_ptr()), C.c_void_p(steps.data_ptr()),
                                     C.c_void_p(status.data_ptr()), None)
    return st, fx.cpu().numpy(), steps.cpu().numpy(), status.cpu().numpy()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("net_fn,tf", [(lambda: models.toy_network(48, 120, seed=9), 0.05),
                                       (lambda: models.colloidal(60, seed=8), 0.02)])
def test_noalloc_matches_simulate(cuda, prec, net_fn, tf):
    m = T.Model.from_network(net_fn())
    track = np.array([0, 2, 5], np.int32)
    for R in (3000, 700):  # the second call is smaller than the first
        ref = m.simulate(tf, R, seed=77, precision=prec, track=track, mode="persistent")
        st, fx, steps, status = _noalloc(m, tf, R, track, seed=77, precision=prec)
        assert st == 0, T.lib().tl_last_error()
        assert ref.n_failed == 0
        assert np.array_equal(fx, ref.final_x)
        assert np.array_equal(steps, ref.steps)
        assert np.array_equal(status, ref.status)


def test_noalloc_needs_resident_model(cuda):
    m = T.Model.from_network(models.toy_network(20, 40, seed=3))
    st, _, _, _ = _noalloc(m, 0.01, 64, np.array([0], np.int32))
    assert st == 9
    assert b"not on device" in T.lib().tl_last_error()
    m.simulate(0.01, 64, track=[0])
    st, fx, _, _ = _noalloc(m, 0.01, 64, np.array([0], np.int32))
    assert st == 0 and (fx >= 0).all()
