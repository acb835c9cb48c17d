"""fw_wigner_propagate_noalloc (the entry point that allocates nothing: its
run scratch is a static device array) against fw_wigner_propagate, through the
C ABI, on two calls in a row, the second on a smaller grid; and its refusals."""
import ctypes

import numpy as np
import pytest

from fastwigner_amd import _lib

pytestmark = pytest.mark.gpu

PHYS = dict(X_amplitude=8., P_amplitude=8., dt=0.01, t=0., V="0.5*X*X+0.1*X*X*X*X", K="0.5*P*P",
            diff_V="X+0.4*X*X*X", diff_K="P")


def _make(n, **kw):
    from fastwigner_amd.wigner_moyal import WignerMoyalHIP1D
    s = WignerMoyalHIP1D(X_gridDIM=n, P_gridDIM=n, **PHYS, **kw)
    s.set_wignerfunction(np.exp(-(s.X - 1.) ** 2 - (s.P - 0.5) ** 2) / np.pi)
    return s


def _state_ptr(s):
    p, es = ctypes.c_void_p(), ctypes.c_int()
    _lib.call("fw_wigner_state_ptr", s._handle, ctypes.byref(p), ctypes.byref(es))
    return p


def _compare(n, steps):
    ref, runner, buf = _make(n), _make(n), _make(n)
    w_runner = runner.wignerfunction.get()
    times = np.array([(s + 1) * PHYS["dt"] for s in range(steps)])
    ser_ref = np.empty((steps, _lib.FW_NSER))
    ser_na = np.empty((steps, _lib.FW_NSER))
    _lib.call("fw_wigner_propagate", ref._handle, _lib.dptr(times), steps, _lib.dptr(ser_ref))
    # the no-allocation run on the W buffer of a third handle (any caller-owned device buffer)
    _lib.call("fw_wigner_propagate_noalloc", runner._handle, _state_ptr(buf), _lib.dptr(times), steps,
              _lib.dptr(ser_na))
    a, b = ref.wignerfunction.get(), buf.wignerfunction.get()
    err = np.linalg.norm(b - a) / np.linalg.norm(a)
    print("noalloc %dx%d, %d steps: rel L2 vs fw_wigner_propagate %.2e, series max diff %.2e"
          % (n, n, steps, err, np.max(np.abs(ser_na - ser_ref))), flush=True)
    assert err <= 1e-12
    assert np.max(np.abs(ser_na - ser_ref)) <= 1e-12 * np.max(np.abs(ser_ref))
    # the runner handle's own state is untouched
    assert np.array_equal(runner.wignerfunction.get(), w_runner)


def test_noalloc_matches_propagate_twice(gpu):
    """1024^2 over 100 steps (a whole chunk and a partial one), then a smaller
    256^2 grid over 20 steps: the static scratch keeps the first call's
    contents, which the second must not read."""
    _compare(1024, 100)
    _compare(256, 20)


def test_noalloc_refusals(gpu):
    s = _make(64, step_passes=3)
    times = np.array([0.01])
    with pytest.raises(_lib.FastWignerError, match="two-pass"):
        _lib.call("fw_wigner_propagate_noalloc", s._handle, _state_ptr(s), _lib.dptr(times), 1, None)
    # 4096 x 16 with one column pair per block: 2048 blocks per column pass
    big = _make_rect(4096, 16, tiles=(8, 1, 16, 1))
    with pytest.raises(_lib.FastWignerError, match="at most 1024"):
        _lib.call("fw_wigner_propagate_noalloc", big._handle, _state_ptr(big), _lib.dptr(times), 1, None)


def _make_rect(nx, npp, **kw):
    from fastwigner_amd.wigner_moyal import WignerMoyalHIP1D
    return WignerMoyalHIP1D(X_gridDIM=nx, P_gridDIM=npp, **PHYS, **kw)
