"""ctypes binding of libfastwigner.so (C ABI: include/fastwigner.h).

Follows the reference's own FFI pattern for its native crossing, cufft.py:16-29
(ctypes library handle, argtypes/restype per entry point) and cufft.py:31-102
(non-zero status codes raised as typed exceptions).

The library is built in-tree by ``python -m fastwigner_amd.build``. There is no
fallback: if the library is missing every propagator raises ``FastWignerError``.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfastwigner.so")
KCACHE = os.path.join(HERE, "_kcache")

FW_NSER = 9


class FastWignerError(RuntimeError):
    """Base class of the library errors (cf. cuFFT's cufftError, cufft.py:31)."""


class InvalidArgumentError(FastWignerError):
    pass


class CompileError(FastWignerError):
    """hipRTC rejected the generated source, e.g. a malformed V/K/mask string."""


class HipError(FastWignerError):
    pass


class NoDeviceError(FastWignerError):
    pass


class UnsupportedError(FastWignerError):
    pass


_ERRORS = {-1: InvalidArgumentError, -2: CompileError, -3: HipError, -4: NoDeviceError, -5: UnsupportedError}


class WignerDesc(ctypes.Structure):
    _fields_ = [
        ("x_grid_dim", ctypes.c_int), ("p_grid_dim", ctypes.c_int),
        ("x_amplitude", ctypes.c_double), ("p_amplitude", ctypes.c_double),
        ("dt", ctypes.c_double), ("phase_time_shift", ctypes.c_double),
        ("precision", ctypes.c_int), ("device", ctypes.c_int),
        ("col_e", ctypes.c_int), ("col_p", ctypes.c_int),
        ("row_e", ctypes.c_int), ("row_p", ctypes.c_int),
        ("phase_tables", ctypes.c_int), ("step_passes", ctypes.c_int),
        ("v_table", ctypes.c_int),
    ]


class WignerExprs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_char_p) for n in (
        "consts", "V", "K", "diff_V", "diff_K",
        "abs_boundary_x_theta", "abs_boundary_lambda_p", "abs_boundary_x_p")]


class RhoDesc(ctypes.Structure):
    _fields_ = [
        ("x_grid_dim", ctypes.c_int), ("precision", ctypes.c_int),
        ("x_amplitude", ctypes.c_double), ("dt", ctypes.c_double),
        ("phase_time_shift", ctypes.c_double), ("device", ctypes.c_int),
        ("row_e", ctypes.c_int), ("row_p", ctypes.c_int), ("col_e", ctypes.c_int), ("col_p", ctypes.c_int),
    ]


class RhoExprs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_char_p) for n in (
        "consts", "V", "K", "diff_V", "diff_K", "abs_boundary_x", "abs_boundary_p")]


FW_BATCH_NPRM = 8


class BatchDesc(ctypes.Structure):
    _fields_ = [
        ("n_instances", ctypes.c_int), ("precision", ctypes.c_int), ("device", ctypes.c_int),
        ("mode", ctypes.c_int), ("dt", ctypes.c_double), ("phase_time_shift", ctypes.c_double),
        ("e_for_n", ctypes.c_int * 5), ("blocks_per_cu", ctypes.c_int),
    ]


class BatchInstance(ctypes.Structure):
    _fields_ = [
        ("grid_dim", ctypes.c_int), ("x_amplitude", ctypes.c_double), ("p_amplitude", ctypes.c_double),
        ("params", ctypes.c_double * FW_BATCH_NPRM),
    ]


class BatchExprs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_char_p) for n in ("consts", "param_names", "V", "K", "diff_V", "diff_K")]


class FnoDesc(ctypes.Structure):
    _fields_ = [("grid_dim", ctypes.c_int), ("width", ctypes.c_int), ("modes", ctypes.c_int),
                ("layers", ctypes.c_int), ("max_batch", ctypes.c_int), ("device", ctypes.c_int),
                ("x_amplitude", ctypes.c_double), ("p_amplitude", ctypes.c_double)]


def uses_token(expr, tok):
    """Whether the C expression references the identifier token ``tok`` (the
    library's own test for a time-dependent string, fw_wigner.cpp uses_token)."""
    import re
    return bool(expr) and re.search(r"(?<![A-Za-z0-9_])%s(?![A-Za-z0-9_])" % re.escape(tok), expr) is not None


def _b(s):
    return None if s is None else s.encode()


def wigner_desc(nx, xamp, npp, pamp, dt, shift=0.5, precision=0, device=0, tiles=None, phase_tables=0,
                step_passes=0, v_table=0):
    d = WignerDesc()
    d.x_grid_dim, d.p_grid_dim = int(nx), int(npp)
    d.x_amplitude, d.p_amplitude = float(xamp), float(pamp)
    d.dt, d.phase_time_shift = float(dt), float(shift)
    d.precision, d.device = int(precision), int(device)
    if tiles:
        d.col_e, d.col_p, d.row_e, d.row_p = [int(v) for v in tiles]
    d.phase_tables = int(phase_tables)
    d.step_passes = int(step_passes)
    d.v_table = int(v_table)
    return d


def wigner_exprs(V, K, consts=None, diff_V=None, diff_K=None, abs_boundary_x_theta=None,
                 abs_boundary_lambda_p=None, abs_boundary_x_p=None):
    e = WignerExprs()
    e.consts, e.V, e.K = _b(consts), _b(V), _b(K)
    e.diff_V, e.diff_K = _b(diff_V), _b(diff_K)
    e.abs_boundary_x_theta = _b(abs_boundary_x_theta)
    e.abs_boundary_lambda_p = _b(abs_boundary_lambda_p)
    e.abs_boundary_x_p = _b(abs_boundary_x_p)
    return e


FW_RSER = 20


def rho_desc(n, xamp, dt, shift=0.0, precision=0, device=0, tiles=None):
    d = RhoDesc()
    d.x_grid_dim, d.x_amplitude, d.dt, d.phase_time_shift = int(n), float(xamp), float(dt), float(shift)
    d.precision, d.device = int(precision), int(device)
    if tiles:
        d.row_e, d.row_p, d.col_e, d.col_p = [int(v) for v in tiles]
    return d


def rho_exprs(V, K, consts=None, diff_V=None, diff_K=None, abs_boundary_x=None, abs_boundary_p=None):
    e = RhoExprs()
    e.consts, e.V, e.K, e.diff_V, e.diff_K = _b(consts), _b(V), _b(K), _b(diff_V), _b(diff_K)
    e.abs_boundary_x, e.abs_boundary_p = _b(abs_boundary_x), _b(abs_boundary_p)
    return e


def batch_desc(n, dt, precision=1, device=0, mode=0, shift=0.0, e_for_n=None, blocks_per_cu=0):
    d = BatchDesc()
    d.n_instances, d.precision, d.device, d.mode = int(n), int(precision), int(device), int(mode)
    d.dt, d.phase_time_shift = float(dt), float(shift)
    for k, e in enumerate(e_for_n or [0] * 5):
        d.e_for_n[k] = int(e)
    d.blocks_per_cu = int(blocks_per_cu)
    return d


def batch_instances(specs):
    """specs: list of (grid_dim, x_amplitude, p_amplitude, params)."""
    arr = (BatchInstance * len(specs))()
    for i, (n, xa, pa, prm) in enumerate(specs):
        arr[i].grid_dim, arr[i].x_amplitude, arr[i].p_amplitude = int(n), float(xa), float(pa)
        prm = list(prm)
        if len(prm) > FW_BATCH_NPRM:
            raise ValueError("at most %d parameters per instance" % FW_BATCH_NPRM)
        for k, v in enumerate(prm):
            arr[i].params[k] = float(v)
    return arr


def batch_exprs(V, K, param_names="", consts=None, diff_V=None, diff_K=None):
    e = BatchExprs()
    e.consts, e.param_names, e.V, e.K = _b(consts), _b(param_names), _b(V), _b(K)
    e.diff_V, e.diff_K = _b(diff_V), _b(diff_K)
    return e


_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_PD = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)
_PP = ctypes.POINTER(ctypes.c_void_p)
_PF = ctypes.POINTER(ctypes.c_float)

# name -> (argtypes); every entry point returns int
PROTOTYPES = {
    "fw_version": [],
    "fw_device_count": [_PI],
    "fw_wigner_grid_constants": [ctypes.POINTER(WignerDesc), _PD],
    "fw_wigner_check": [ctypes.POINTER(WignerDesc), ctypes.POINTER(WignerExprs)],
    "fw_wigner_prelude": [ctypes.POINTER(WignerDesc), ctypes.POINTER(WignerExprs), ctypes.c_char_p, ctypes.c_size_t,
                          ctypes.POINTER(ctypes.c_size_t)],
    "fw_wigner_create": [ctypes.POINTER(WignerDesc), ctypes.POINTER(WignerExprs), _PP],
    "fw_wigner_destroy": [_P],
    "fw_wigner_set_state": [_P, _PD],
    "fw_wigner_set_state_expr": [_P, ctypes.c_char_p],
    "fw_wigner_fill": [_P, _D],
    "fw_wigner_normalize": [_P, _PD],
    "fw_wigner_get_state": [_P, _PD],
    "fw_wigner_state_ptr": [_P, _PP, _PI],
    "fw_wigner_propagate": [_P, _PD, _I, _PD],
    "fw_wigner_single_step": [_P, _D, _PD],
    "fw_wigner_propagate_noalloc": [_P, _P, _PD, _I, _PD],
    "fw_wigner_moments": [_P, _D, _PD],
    "fw_wigner_get_average": [_P, ctypes.c_char_p, _D, _PD],
    "fw_wigner_bloch_setup": [_P, _I, _D, _PD],
    "fw_wigner_bloch_propagate": [_P, _D, _I, _D, _PD, _PI],
    "fw_wigner_set_dissipator": [_P, _PD, _D],
    "fw_wigner_snapshot": [_P, _I],
    "fw_wigner_set_timing": [_P, _I],
    "fw_wigner_kernel_times": [_P, _PD],
    "fw_wigner_kernel_bench": [_P, ctypes.c_double, _I, _PD],
    "fw_wigner_step_passes": [_P, _PI],
    "fw_wigner_k_factors": [_P, _PI, _PD],
    "fw_rho_grid_constants": [ctypes.POINTER(RhoDesc), _PD],
    "fw_rho_check": [ctypes.POINTER(RhoDesc), ctypes.POINTER(RhoExprs)],
    "fw_rho_create": [ctypes.POINTER(RhoDesc), ctypes.POINTER(RhoExprs), _PP],
    "fw_rho_destroy": [_P],
    "fw_rho_set_state": [_P, _PD],
    "fw_rho_set_state_expr": [_P, ctypes.c_char_p],
    "fw_rho_normalize": [_P, _PD],
    "fw_rho_get_state": [_P, _PD],
    "fw_rho_state_ptr": [_P, _PP, _PI],
    "fw_rho_propagate": [_P, _PD, _I, _I, _PD],
    "fw_rho_single_step": [_P, _D, _PD],
    "fw_rho_moments": [_P, _D, _PD],
    "fw_rho_get_average": [_P, ctypes.POINTER(ctypes.c_char_p), _I, _D, _PD],
    "fw_rho_bloch_setup": [_P, _D, _PD],
    "fw_rho_bloch_propagate": [_P, _D, _I, _D, _PD, _PI],
    "fw_rho_snapshot": [_P, _I],
    "fw_rho_wigner": [_P, _PD],
    "fw_rho_set_wigner_initial": [_P, _PD],
    "fw_rho_wigner_time": [_P, _D, _D, _PD],
    "fw_rho_propagate_wigner_time": [_P, _PD, _I, _I, _D, _PD, _PD],
    "fw_rho_set_timing": [_P, _I],
    "fw_rho_kernel_times": [_P, _PD],
    "fw_batch_check": [ctypes.POINTER(BatchDesc), ctypes.POINTER(BatchInstance), ctypes.POINTER(BatchExprs)],
    "fw_batch_create": [ctypes.POINTER(BatchDesc), ctypes.POINTER(BatchInstance), ctypes.POINTER(BatchExprs), _PP],
    "fw_batch_destroy": [_P],
    "fw_batch_set_state": [_P, _I, _PD],
    "fw_batch_get_state": [_P, _I, _PD],
    "fw_batch_set_dissipator": [_P, _I, _PD, _D],
    "fw_batch_gibbs_dissipators": [_P, _PD],
    "fw_batch_bloch_propagate": [_P, _I, _D, _PD, _PI],
    "fw_batch_propagate": [_P, _PD, _I, _PD],
    "fw_batch_info": [_P, ctypes.POINTER(ctypes.c_longlong), _PI, _PI],
    "fw_batch_step_time": [_P, _I, _PD],
    "fw_batch_state_ptr": [_P, _I, _PP, _PI],
    "fw_batch_copy_states": [_P, _P],
    "fw_batch_passes": [_P, _PI],
    "fw_fno_create": [ctypes.POINTER(FnoDesc), _PP],
    "fw_fno_destroy": [_P],
    "fw_fno_param_count": [_P, ctypes.POINTER(ctypes.c_longlong)],
    "fw_fno_set_params": [_P, _PF],
    "fw_fno_get_params": [_P, _PF],
    "fw_fno_set_output": [_P, ctypes.c_float, _I],
    "fw_fno_forward": [_P, _PF, _PF, _I, _PF],
    "fw_fno_moments": [_P, _I, _PD],
    "fw_fno_loss_grad": [_P, _PF, _PF, _PF, _I, _PD, _PF],
    "fw_fno_dataset_reserve": [_P, _I],
    "fw_fno_dataset_put": [_P, _I, _P, _P, _PF],
    "fw_fno_train_step": [_P, _PI, _I, ctypes.c_float, _PD],
    "fw_fno_eval_pairs": [_P, _PI, _I, _PD],
    "fw_fno_infer_time": [_P, _I, _I, _PD],
}

_lib = None


def load(force_reload=False):
    """Load libfastwigner.so (no fallback) and attach prototypes."""
    global _lib
    if _lib is not None and not force_reload:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FastWignerError(
            "libfastwigner.so not found at %s; build it with `python -m fastwigner_amd.build`" % LIB_PATH)
    if "FASTWIGNER_CACHE_DIR" not in os.environ and os.access(HERE, os.W_OK):
        os.environ["FASTWIGNER_CACHE_DIR"] = KCACHE
    lib = ctypes.CDLL(LIB_PATH)
    lib.fw_last_error.restype = ctypes.c_char_p
    lib.fw_last_error.argtypes = []
    for name, args in PROTOTYPES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _lib = lib
    return lib


def check(rc, what=""):
    """Raise the typed exception of a non-zero status (cufft.py:82-102)."""
    if rc != 0:
        msg = load().fw_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, FastWignerError)("%s: %s" % (what, msg) if what else msg)
    return rc


def call(name, *args):
    return check(getattr(load(), name)(*args), name)


def dptr(a):
    """ctypes double* to a contiguous float64 numpy array."""
    return a.ctypes.data_as(_PD)


def fptr(a):
    """ctypes float* to a contiguous float32 numpy array."""
    assert a.dtype == "float32" and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_PF)
