"""Host-only checks of the C ABI (no GPU needed): the library loads, exports every
entry point include/*.h declares, validates arguments, maps status codes to typed
exceptions, and hipRTC compiles the device code for gfx950 (or reports the
compiler log for a malformed expression)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from fastwigner_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(fw_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_device_count():
    lib = _lib.load()
    assert lib.fw_version() >= 100
    n = ctypes.c_int(-1)
    assert lib.fw_device_count(ctypes.byref(n)) == 0
    assert n.value >= 0


@pytest.mark.parametrize("nx,npp,msg", [(63, 64, "even"), (4, 64, "in [8, 4096]"), (8192, 64, "in [8, 4096]"),
                                        (34, 64, "at most 15"), (2 * 3 * 7 * 4, 64, "at most 15")])
def test_invalid_grids(nx, npp, msg):
    lib = _lib.load()
    d = _lib.wigner_desc(nx, 8., npp, 8., 0.01)
    e = _lib.wigner_exprs(V="0.5*X*X", K="0.5*P*P")
    rc = lib.fw_wigner_check(ctypes.byref(d), ctypes.byref(e))
    assert rc == -1
    assert msg in lib.fw_last_error().decode()
    with pytest.raises(_lib.InvalidArgumentError):
        _lib.check(rc)


def test_invalid_tiles_and_null_args():
    lib = _lib.load()
    d = _lib.wigner_desc(64, 8., 64, 8., 0.01, tiles=(3, 1, 8, 1))
    e = _lib.wigner_exprs(V="0.5*X*X", K="0.5*P*P")
    assert lib.fw_wigner_check(ctypes.byref(d), ctypes.byref(e)) == -1
    assert lib.fw_wigner_check(None, ctypes.byref(e)) == -1
    assert lib.fw_wigner_destroy(None) == 0
    assert lib.fw_wigner_propagate(None, None, 1, None) == -1
    gen, dev = ctypes.c_int(), ctypes.c_double()
    assert lib.fw_wigner_k_factors(None, ctypes.byref(gen), ctypes.byref(dev)) == -1


@pytest.mark.parametrize("n,tiles,ok", [
    (1000, (10, 1, 20, 1), True),    # 2^3 5^3: radix-10 and radix-20 stages
    (1000, (8, 1, 8, 1), False),     # E without the factor 5
    (768, (12, 1, 24, 2), True),     # 2^8 3
    (768, (16, 1, 16, 1), False),    # E a power of two on a line with a factor 3
    (120, (30, 1, 30, 1), True),     # 2^3 3 5: E holds both odd primes
    (120, (15, 1, 30, 1), False),    # odd E on a line that is not a power of two
    (64, (6, 1, 8, 1), False),       # power-of-two line: power-of-two E only
])
def test_even_grid_tile_rules(n, tiles, ok):
    """The schedules of lines that are not powers of two: E divides the
    line, is even and holds a radix for every stage (checked before any
    compilation); a valid schedule is only checked by the GPU tests."""
    lib = _lib.load()
    d = _lib.wigner_desc(n, 8., n, 8., 0.01, tiles=tiles)
    e = _lib.wigner_exprs(V="0.5*X*X", K="0.5*P*P")
    plen = ctypes.c_size_t()
    rc = lib.fw_wigner_prelude(ctypes.byref(d), ctypes.byref(e), None, 0, ctypes.byref(plen))
    assert (rc == 0) == ok, lib.fw_last_error().decode()
    if not ok:
        assert "elements per thread" in lib.fw_last_error().decode()


@pytest.mark.parametrize("n,tiles,msg", [
    (34, None, "at most 15"),            # 2 * 17
    (63, None, "even"),
    (1000, (8, 1, 16, 1), "elements per thread"),   # E without the factor 5
])
def test_density_matrix_size_rules(n, tiles, msg):
    """The density matrix takes the even sizes the Wigner classes take; refused
    sizes and schedules fail before any compilation."""
    lib = _lib.load()
    d = _lib.rho_desc(n, 10., 0.01, tiles=tiles)
    e = _lib.rho_exprs(V="0.5*X*X", K="0.5*P*P")
    assert lib.fw_rho_check(ctypes.byref(d), ctypes.byref(e)) == -1
    assert msg in lib.fw_last_error().decode()


def test_device_code_compiles_for_the_longest_schedules():
    """4096-point lines at four elements per thread: the most stages per
    direction (six, with their base twiddles computed ahead)."""
    lib = _lib.load()
    d = _lib.wigner_desc(4096, 8., 4096, 8., 0.01, tiles=(4, 1, 4, 1))
    e = _lib.wigner_exprs(V="0.5*X*X", K="0.5*P*P")
    rc = lib.fw_wigner_check(ctypes.byref(d), ctypes.byref(e))
    assert rc == 0, lib.fw_last_error().decode()


def test_device_code_compiles_for_an_even_grid():
    """A 30 x 22 grid (one radix-30 and one radix-22 stage per line, prime-
    factor and odd-radix kernels) compiles for gfx950 in both precisions."""
    lib = _lib.load()
    for prec in (0, 1):
        d = _lib.wigner_desc(30, 8., 22, 8., 0.01, precision=prec)
        e = _lib.wigner_exprs(V="0.5*X*X+0.1*X*X*X*X", K="0.5*P*P", diff_V="X+0.4*X*X*X", diff_K="P")
        rc = lib.fw_wigner_check(ctypes.byref(d), ctypes.byref(e))
        assert rc == 0, lib.fw_last_error().decode()


def test_malformed_expression_is_a_compile_error():
    lib = _lib.load()
    d = _lib.wigner_desc(16, 8., 16, 8., 0.01)
    e = _lib.wigner_exprs(V="0.5*X*X +* undefined_thing(", K="0.5*P*P")
    rc = lib.fw_wigner_check(ctypes.byref(d), ctypes.byref(e))
    assert rc == -2
    log = lib.fw_last_error().decode()
    assert "undefined_thing" in log
    with pytest.raises(_lib.CompileError):
        _lib.check(rc)


def test_device_code_compiles_for_gfx950():
    """The fused step kernels, Bloch and dissipator sets compile for gfx950 with
    constants and functions in the reference's style (strong-field E(t))."""
    lib = _lib.load()
    consts = ("    const double F = 0.04;\n    const double omega = 0.05698;\n"
              "__device__ double E(double t)\n{{\n  return -F * sin(omega * t) * pow(sin(omega * t / 16.), 2);\n}}\n")
    d = _lib.wigner_desc(32, 60., 16, 6., 0.05, precision=1)
    e = _lib.wigner_exprs(V="-1. / sqrt(X * X + 1.) + X * E(t)", K="0.5 * P * P", consts=consts,
                          diff_V="X / pow(X * X + 1., 1.5) + E(t)", diff_K="P",
                          abs_boundary_x_p="pow(fabs(cos(0.5 * M_PI * X / X_amplitude)), 0.125)")
    rc = lib.fw_wigner_check(ctypes.byref(d), ctypes.byref(e))
    assert rc == 0, lib.fw_last_error().decode()


def test_grid_constants_match_numpy():
    import numpy as np
    d = _lib.wigner_desc(1024, 60., 512, 6., 0.05)
    out = (ctypes.c_double * 4)()
    assert _lib.load().fw_wigner_grid_constants(ctypes.byref(d), out) == 0
    dX = 2. * 60. / 1024
    dP = 2. * 6. / 512
    lam = np.fft.fftfreq(1024, dX / (2 * np.pi))
    th = np.fft.fftfreq(512, dP / (2 * np.pi))
    assert list(out) == [dX, dP, lam[1] - lam[0], th[1] - th[0]]


def test_product_fails_loudly_without_gpu():
    n = ctypes.c_int(0)
    _lib.load().fw_device_count(ctypes.byref(n))
    if n.value:
        pytest.skip("a GPU is visible")
    from fastwigner_amd.wigner_moyal import WignerMoyalHIP1D
    with pytest.raises(_lib.NoDeviceError):
        WignerMoyalHIP1D(X_gridDIM=16, X_amplitude=8., P_gridDIM=16, P_amplitude=8., dt=0.01, t=0.,
                         V="0.5*X*X", K="0.5*P*P")


def test_rho_record_matches_per_step_formula():
    """The vectorised Ehrenfest recording of the density-matrix class equals its
    per-step formula (normalised traces of the raw sums) on random sums."""
    from fastwigner_amd.rho_vneumann import RhoVNeumannHIP1D

    class Holder:
        dX = 0.037

    h = Holder()
    names = ("X_average", "X_average_RHS", "P_average", "P_average_RHS", "hamiltonian_average", "purity_series")
    for k in names:
        setattr(h, k, [])
    ser = np.random.default_rng(0).standard_normal((7, 20))
    RhoVNeumannHIP1D._record(h, ser)
    for i, row in enumerate(ser):
        o = RhoVNeumannHIP1D._observables(h, row)
        for k in names:
            ref = o["purity" if k == "purity_series" else k]
            assert abs(getattr(h, k)[i] - ref) <= 1e-14 * max(1.0, abs(ref)), k


def _prelude(V, consts=None, diff_V=None, **kw):
    d = _lib.wigner_desc(64, 8., 64, 8., 0.01, 0.5, 0, 0, None, **kw)
    e = _lib.wigner_exprs(V=V, K="0.5*P*P", consts=consts, diff_V=diff_V, diff_K="P" if diff_V else None)
    n = ctypes.c_size_t()
    _lib.call("fw_wigner_prelude", ctypes.byref(d), ctypes.byref(e), None, 0, ctypes.byref(n))
    buf = ctypes.create_string_buffer(n.value)
    _lib.call("fw_wigner_prelude", ctypes.byref(d), ctypes.byref(e), buf, n.value, ctypes.byref(n))
    return buf.value.decode()


@pytest.mark.parametrize("V,split", [
    ("-1. / sqrt(X * X + 1.) + X * E(t)", True), ("0.5*X*X + X*E0*sin(om*t)", True),
    ("E(t)*X - 0.5*X", True), ("1e-5*X - X*sin(t)", True),
    ("0.5*X*X*(1. + 0.05*sin(0.3*t))", False), ("X*X*E(t)", False), ("0.5*X*X", False),
    ("X*sin(t)/2", False), ("-X*sin(t)", False), ("X*sin(t) + E(t)", False),
    # operators that bind more loosely than + and - at the top level: never split
    ("X > 0. ? 0.5*X*X : 0.1*X*X + X*E(t)", False), ("(X > 0.) + X*E(t)", True),
    ("X < 1. || t > 2. ? 0.5*X*X + X*E(t) : 0.", False), ("0.5*X*X + X*E(t) == 0.", False),
    ("0.5*X*X + X*E(t), 0.", False), ("0.5*X*X + X*E(t) && 1", False)])
def test_field_split_of_time_dependent_potentials(V, split):
    """V = V0(X) + X F(t) at the top level of the expression gets a table of V0
    and a per-step field factor (FW_V_SPLIT); every other time-dependent form
    (and a time-independent V) is left whole. The split parts, evaluated with
    numpy, add up to V."""
    consts = "__device__ double E(double t) { return 0.05*sin(0.06*t); }\nconst double E0 = 0.05; const double om = 0.3;"
    src = _prelude(V, consts)
    assert ("#define FW_V_SPLIT 1" in src) == split
    if not split:
        return
    vs = re.search(r"fw_V_static\(double X, double t\) \{ \(void\)t; return \((.*)\); \}", src).group(1)
    vf = re.search(r"fw_V_field\(double t\) \{ return \((.*)\); \}", src).group(1)
    env = dict(sqrt=np.sqrt, sin=np.sin, E=lambda t: 0.05 * np.sin(0.06 * t), E0=0.05, om=0.3)
    for X in (-3.0, 0.0, 0.7, 5.0):
        for t in (0.0, 1.3, 40.0):
            full = eval(V, dict(env, X=X, t=t))
            parts = eval(vs, dict(env, X=X, t=t)) + X * eval(vf, dict(env, t=t))
            assert abs(full - parts) <= 1e-12 * max(1.0, abs(full))


def test_field_split_needs_tables():
    """With phase_tables = -1 (factors evaluated at every half-step) V is not split."""
    src = _prelude("0.5*X*X + X*E0*sin(om*t)", "const double E0 = 0.05; const double om = 0.3;", phase_tables=-1)
    assert "#define FW_V_SPLIT 1" not in src


@pytest.mark.parametrize("dV,split", [("X / pow(X * X + 1., 1.5) + E(t)", True), ("X - 2.*E0*sin(om*t)", True),
                                      ("X*(1. + E(t))", False), ("X + X*E(t)", False), ("X", False),
                                      ("X > 0. ? X : 0.2*X + E(t)", False), ("X != 0. ? X + E(t) : E(t)", False)])
def test_field_split_of_the_force(dV, split):
    """With a split V, V' = dV0(X) + G(t) (time-dependent terms free of X) is
    split too (FW_DV_SPLIT) and the parts add up to V'."""
    consts = "__device__ double E(double t) { return 0.05*sin(0.06*t); }\nconst double E0 = 0.05; const double om = 0.3;"
    src = _prelude("-1. / sqrt(X * X + 1.) + X * E(t)", consts, diff_V=dV)
    assert "#define FW_V_SPLIT 1" in src
    assert ("#define FW_DV_SPLIT 1" in src) == split
    if not split:
        return
    ds = re.search(r"fw_dV_static\(double X, double t\) \{ \(void\)t; return \((.*)\); \}", src).group(1)
    dg = re.search(r"fw_dV_field\(double t\) \{ return \((.*)\); \}", src).group(1)
    env = dict(sqrt=np.sqrt, sin=np.sin, pow=np.power, E=lambda t: 0.05 * np.sin(0.06 * t), E0=0.05, om=0.3)
    for X in (-3.0, 0.0, 0.7, 5.0):
        for t in (0.0, 1.3, 40.0):
            full = eval(dV, dict(env, X=X, t=t))
            parts = eval(ds, dict(env, X=X, t=t)) + eval(dg, dict(env, t=t))
            assert abs(full - parts) <= 1e-12 * max(1.0, abs(full))
