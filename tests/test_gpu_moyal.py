"""GPU parity of the plain Moyal propagator against the CPU oracle.

Tolerances: FP64 relative L2 <= 1e-12 against the oracle (the parity bar; the
FFT-to-FFT noise floor measured between pocketfft and torch on the quartic 256^2 workload is
1.6e-13). FP32 form: relative L2 <= 1e-4 against the FP64 oracle.
"""
import numpy as np
import pytest

from oracle import cexpr
from oracle import fw_oracle as fo

pytestmark = pytest.mark.gpu

V_Q = "0.5*X*X+0.1*X*X*X*X"
K_Q = "0.5*P*P"
DV_Q = "X+0.4*X*X*X"
DK_Q = "P"


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def make_pair(nx, xa, npp, pa, dt, V, K, dV=None, dK=None, consts=None, masks=None, shift=0.5,
              precision="fp64", tiles=None, **kw):
    from fastwigner_amd.wigner_moyal import WignerMoyalHIP1D
    masks = masks or {}
    cvals = dict(consts or {})
    cx = dict(cvals, X_gridDIM=nx, P_gridDIM=npp, X_amplitude=xa, P_amplitude=pa, dt=dt,
              dX=2. * xa / nx, dP=2. * pa / npp)
    args = dict(X_gridDIM=nx, X_amplitude=xa, P_gridDIM=npp, P_amplitude=pa, dt=dt, t=0.0, V=V, K=K,
                precision=precision, phase_time_shift=shift, **cvals, **masks, **kw)
    if dV:
        args.update(diff_V=dV, diff_K=dK)
    if tiles:
        args["tiles"] = tiles
    g = WignerMoyalHIP1D(**args)
    o = fo.MoyalOracle(
        nx, xa, npp, pa, dt, V=cexpr.make(V, cx), K=cexpr.make(K, cx, ("P", "t")),
        diff_V=cexpr.make(dV, cx) if dV else None, diff_K=cexpr.make(dK, cx, ("P", "t")) if dK else None,
        phase_time_shift=shift,
        mask_x_theta=cexpr.make(masks.get("abs_boundary_x_theta"), cx, ("X", "Theta", "t"), cexpr.x_theta_names),
        mask_lambda_p=cexpr.make(masks.get("abs_boundary_lambda_p"), cx, ("Lambda", "P", "t")),
        mask_x_p=cexpr.make(masks.get("abs_boundary_x_p"), cx, ("X", "P", "t")))
    return g, o


def gaussian(o, x0=1.0, p0=0.5):
    return np.exp(-(o.X - x0) ** 2 - (o.P - p0) ** 2) / np.pi


def k_factors(g):
    """(generated, max |generated - table|) of fw_wigner_k_factors."""
    import ctypes
    from fastwigner_amd import _lib
    gen, dev = ctypes.c_int(), ctypes.c_double()
    _lib.call("fw_wigner_k_factors", g._handle, ctypes.byref(gen), ctypes.byref(dev))
    return gen.value, dev.value


def test_generated_k_factors(gpu):
    """K = P^2/2 in FP64: the K passes form their factors as powers of one phase per row
    (checked against the table at creation) and 500 steps match the oracle at the parity
    bar and the table-reading run (phase_tables = 1) to 1e-13, both passes per step
    forms; a K with a P^4 term, a lambda-p mask and the FP32 form read the table."""
    for passes in (0, 3):
        g, o = make_pair(512, 8., 512, 8., 0.01, V_Q, K_Q, DV_Q, DK_Q, step_passes=passes)
        gt, _ = make_pair(512, 8., 512, 8., 0.01, V_Q, K_Q, DV_Q, DK_Q, step_passes=passes, phase_tables=1)
        gen, dev = k_factors(g)
        print("passes %d: generated %d, max |generated - table| %.2e" % (passes or 2, gen, dev))
        assert gen == 1 and 0 <= dev <= 1e-13
        assert k_factors(gt) == (0, -1.0)
        W0 = gaussian(o)
        for s in (g, gt, o):
            s.set_wignerfunction(W0)
        Wg = g.propagate(500).get()
        Wt = gt.propagate(500).get()
        o.propagate(500)
        e_o, e_t = rel_l2(Wg, o.W), rel_l2(Wg, Wt)
        print("  rel L2 against the oracle %.2e, against the table run %.2e" % (e_o, e_t))
        assert e_o <= 1e-12 and e_t <= 1e-13
        assert np.allclose(g.P_average, gt.P_average, rtol=0, atol=1e-13)
    gq, _ = make_pair(64, 8., 64, 8., 0.01, V_Q, "0.5*P*P+0.01*P*P*P*P")
    gen, dev = k_factors(gq)
    assert gen == 0 and dev > 1e-11
    gm, _ = make_pair(64, 8., 64, 8., 0.01, V_Q, K_Q,
                      masks={"abs_boundary_lambda_p": "exp(-0.001*Lambda*Lambda)"})
    gen, dev = k_factors(gm)
    assert gen == 0 and dev > 1e-11
    g32, _ = make_pair(256, 8., 256, 8., 0.01, V_Q, K_Q, precision="fp32")
    assert k_factors(g32) == (0, -1.0)


def test_short_run_graphs_match_launches(gpu):
    """Repeated runs of one shape replay a captured graph from the second call on
    (opening V/2, the steps, the closing normalisation); W and every series value equal
    those of the same calls launched kernel by kernel (timing on: no graphs), bit for
    bit, over short, multi-chunk and one-step runs, and the whole sequence matches the
    oracle stepped through it at the parity bar."""
    from fastwigner_amd import _lib
    g, o = make_pair(256, 8., 256, 8., 0.01, V_Q, K_Q, DV_Q, DK_Q)
    h, _ = make_pair(256, 8., 256, 8., 0.01, V_Q, K_Q, DV_Q, DK_Q)
    _lib.call("fw_wigner_set_timing", h._handle, 1)
    W0 = gaussian(o)
    for s in (g, h, o):
        s.set_wignerfunction(W0)
    for n in (20, 20, 20, 7, 100, 100, 64, 64, 1, 1, 20):
        Wg = g.propagate(n).get()
        Wh = h.propagate(n).get()
        assert np.array_equal(Wg, Wh), n
        o.propagate(n)
    for name in ("X_average", "X_average_RHS", "P_average", "P_average_RHS", "hamiltonian_average",
                 "purity_series"):
        assert getattr(g, name) == getattr(h, name), name
    err = rel_l2(Wg, o.W)
    print("after 417 steps in 11 calls: rel L2 against the oracle %.2e" % err)
    assert err <= 1e-12
    op = np.array([x["P_average"] for x in o.series])
    assert len(g.P_average) == len(op) == 417
    assert np.max(np.abs(np.array(g.P_average) - op)) <= 1e-11


@pytest.mark.parametrize("tables,passes", [(0, 0), (-1, 0), (0, 3), (-1, 3)])
def test_quartic_256_parity_fp64(gpu, tables, passes):
    """Quartic Moyal workload: 256^2, x,p in [-8,8), V=0.5x^2+0.1x^4, dt=0.01, 1000 steps,
    <x>,<p>,<H>, purity logged each step; with cached factor tables (default for
    time-independent V, K) and with the factors evaluated at every half-step;
    in two fused passes per step (default) and in three."""
    g, o = make_pair(256, 8., 256, 8., 0.01, V_Q, K_Q, DV_Q, DK_Q, phase_tables=tables, step_passes=passes)
    W0 = gaussian(o)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    Wg = g.propagate(1000).get()
    o.propagate(1000)
    err = rel_l2(Wg, o.W)
    print("quartic 256^2 fp64 rel L2 vs oracle:", err)
    assert err <= 1e-12
    ox = np.array([s["X_average"] for s in o.series])
    op = np.array([s["P_average"] for s in o.series])
    oh = np.array([s["hamiltonian_average"] for s in o.series])
    opu = np.array([s["purity"] for s in o.series])
    assert np.max(np.abs(np.array(g.X_average) - ox)) <= 1e-11
    assert np.max(np.abs(np.array(g.P_average) - op)) <= 1e-11
    assert np.max(np.abs(np.array(g.hamiltonian_average) - oh)) <= 1e-11 * np.max(np.abs(oh))
    assert np.max(np.abs(np.array(g.purity_series) - opu)) <= 1e-11
    assert abs(g.t - o.t) == 0.0
    from fastwigner_amd import _lib
    import ctypes
    n = ctypes.c_int()
    _lib.call("fw_wigner_step_passes", g._handle, ctypes.byref(n))
    assert n.value == (3 if passes == 3 else 2)


def test_harmonic_closed_form_gpu(gpu):
    """The Strang step is the leapfrog map for V=x^2/2: a Gaussian stays Gaussian
    with mean and covariance propagated by the 2x2 leapfrog matrix."""
    g, o = make_pair(256, 8., 256, 8., 0.01, "0.5*X*X", K_Q)
    g.set_wignerfunction(gaussian(o))
    n = 500
    W = g.propagate(n).get()
    a = 0.5 * 0.01
    M = np.array([[1, 0], [-a, 1]]) @ np.array([[1, 0.01], [0, 1]]) @ np.array([[1, 0], [-a, 1]])
    Mn = np.linalg.matrix_power(M, n)
    mu = Mn @ np.array([1., 0.5])
    S = Mn @ (0.5 * np.eye(2)) @ Mn.T
    Si = np.linalg.inv(S)
    dx, dp = o.X - mu[0], o.P - mu[1]
    Wa = np.exp(-0.5 * (Si[0, 0] * dx * dx + 2 * Si[0, 1] * dx * dp + Si[1, 1] * dp * dp))
    Wa /= Wa.sum() * o.dXdP
    assert rel_l2(W, Wa) <= 1e-12


@pytest.mark.parametrize("nx,npp,xp,passes", [(128, 64, "x", 0), (64, 256, "x", 0), (512, 512, "x", 0),
                                              (128, 64, "xp", 0), (256, 256, "xp", 0), (128, 128, "x", 3)])
def test_rectangular_masks_time_dependent(gpu, nx, npp, xp, passes):
    """Non-square grids, all three absorbing-boundary strings and a
    time-dependent potential V(x, t) = x^2/2 + x E(t) (strong-field pattern;
    a cached table of x^2/2 times a per-step field factor).
    An x-p mask of x alone keeps the fused column pass on the theta spectrum;
    one that also depends on p sends it through x-p space between its halves;
    three passes per step apply the field factor in the plain V pass."""
    consts = {"E0": 0.05, "om": 0.3}
    mxp = "pow(fabs(cos(0.5*M_PI*X/X_amplitude)), 0.125)"
    if xp == "xp":
        mxp += " * exp(-1e-3*P*P*P*P/(P_amplitude*P_amplitude))"
    masks = dict(abs_boundary_x_theta="pow(fabs(cos(0.5*M_PI*X_minus/X_amplitude)*cos(0.5*M_PI*X_plus/X_amplitude)), 0.01)",
                 abs_boundary_lambda_p="exp(-1e-4*Lambda*Lambda)",
                 abs_boundary_x_p=mxp)
    g, o = make_pair(nx, 10., npp, 6., 0.02, "0.5*X*X + X*E0*sin(om*t)", "0.5*P*P",
                     "X + E0*sin(om*t)", "P", consts=consts, masks=masks, step_passes=passes)
    W0 = np.exp(-(o.X + 0.5) ** 2 - (o.P - 0.3) ** 2)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    Wg = g.propagate(200).get()
    o.propagate(200)
    assert rel_l2(Wg, o.W) <= 1e-12
    dip = np.array([s["P_average_RHS"] for s in o.series])
    assert np.max(np.abs(np.array(g.P_average_RHS) - dip)) <= 1e-11


@pytest.mark.parametrize("passes", [0, 3])
def test_single_step_leaves_the_state_unnormalised(gpu, passes):
    """single_step_propagation (wigner_moyal_cuda_1d.py:496-519): one step at the
    current t, t not advanced, nothing recorded, W not renormalised. An x-p mask
    takes about 9% of the norm away in the step, so a renormalised result would
    miss by that much; the propagate after it takes the unnormalised W as it
    is, as the reference's loop does."""
    masks = dict(abs_boundary_x_p="pow(fabs(cos(0.5*M_PI*X/X_amplitude)), 0.5)")
    g, o = make_pair(128, 8., 128, 8., 0.01, V_Q, K_Q, DV_Q, DK_Q, masks=masks, step_passes=passes)
    W0 = gaussian(o, x0=3.0)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    t0 = g.t
    Wg = g.single_step_propagation().get()
    o.single_step_propagation()
    assert g.t == t0 and len(g.X_average) == 0
    assert rel_l2(Wg, o.W) <= 1e-12
    norm = Wg.sum() * o.dXdP
    assert abs(norm - o.W.sum() * o.dXdP) <= 1e-13
    assert norm < 0.95
    Wg = g.propagate(20).get()
    o.propagate(20)
    assert rel_l2(Wg, o.W) <= 1e-12
    ref = np.array([s["X_average"] for s in o.series])
    assert np.max(np.abs(np.array(g.X_average) - ref)) <= 1e-11


@pytest.mark.parametrize("nx,npp,passes", [(64, 32, 0), (64, 32, 3), (16, 16, 0), (32, 64, 0)])
def test_small_grid_observables(gpu, nx, npp, passes):
    """Grids whose column blocks hold fewer threads than there are sums (4 or 2
    threads at N_p = 32 or 16) and fewer than a wave: every logged observable
    against the oracle, in two- and three-pass steps."""
    g, o = make_pair(nx, 8., npp, 8., 0.01, V_Q, K_Q, DV_Q, DK_Q, step_passes=passes)
    W0 = gaussian(o)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    Wg = g.propagate(130).get()
    o.propagate(130)
    assert rel_l2(Wg, o.W) <= 1e-12
    for k, ours in (("X_average", g.X_average), ("P_average", g.P_average), ("X_average_RHS", g.X_average_RHS),
                    ("P_average_RHS", g.P_average_RHS), ("hamiltonian_average", g.hamiltonian_average),
                    ("purity", g.purity_series)):
        ref = np.array([s[k] for s in o.series])
        assert np.max(np.abs(np.array(ours) - ref)) <= 1e-11 * max(1.0, np.max(np.abs(ref))), k


@pytest.mark.parametrize("tables", [0, -1])
def test_time_independent_masks(gpu, tables):
    """All three absorbing-boundary strings with time-independent V and K: the
    masks enter the cached tables (0) and the per-half-step factors (-1) alike."""
    masks = dict(abs_boundary_x_theta="pow(fabs(cos(0.5*M_PI*X_minus/X_amplitude)*cos(0.5*M_PI*X_plus/X_amplitude)), 0.01)",
                 abs_boundary_lambda_p="exp(-1e-4*Lambda*Lambda)",
                 abs_boundary_x_p="pow(fabs(cos(0.5*M_PI*X/X_amplitude)), 0.125)")
    g, o = make_pair(256, 10., 128, 6., 0.02, V_Q, K_Q, DV_Q, DK_Q, masks=masks, phase_tables=tables)
    W0 = np.exp(-(o.X + 0.5) ** 2 - (o.P - 0.3) ** 2)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    Wg = g.propagate(300).get()
    o.propagate(300)
    assert rel_l2(Wg, o.W) <= 1e-12


SERIES = ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average", "purity")


def check_series(g, o, tol=1e-11):
    ours = dict(X_average=g.X_average, P_average=g.P_average, X_average_RHS=g.X_average_RHS,
                P_average_RHS=g.P_average_RHS, hamiltonian_average=g.hamiltonian_average, purity=g.purity_series)
    for k in SERIES:
        ref = np.array([s[k] for s in o.series])
        assert np.max(np.abs(np.array(ours[k]) - ref)) <= tol * max(1.0, np.max(np.abs(ref))), k


@pytest.mark.parametrize("tiles", [(16, 1, 16, 1), (8, 2, 8, 2), (16, 4, 4, 4), (32, 1, 32, 1), (4, 8, 2, 8),
                                   (2, 4, 4, 4)])
def test_tile_schedules_agree(gpu, tiles):
    """Every FFT tile/radix schedule gives the oracle's answer, W and the
    logged sums (the fused column pass takes the sums on the theta spectrum at
    each schedule's own item-to-thread map; column E = 2 uses the separate
    transforms)."""
    g, o = make_pair(256, 8., 256, 8., 0.01, V_Q, K_Q, DV_Q, DK_Q, tiles=tiles)
    W0 = gaussian(o)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    Wg = g.propagate(50).get()
    o.propagate(50)
    assert rel_l2(Wg, o.W) <= 1e-12
    check_series(g, o)


TD_V = {"field": ("0.5*X*X + X*E0*sin(om*t)", "X + E0*sin(om*t)"),
        "scaled": ("0.5*X*X*(1. + E0*sin(om*t))", "X*(1. + E0*sin(om*t))"),
        "static": (V_Q, DV_Q)}


@pytest.mark.parametrize("nx,npp,tables,vkind", [(128, 64, 0, "field"), (64, 256, 0, "field"),
                                                 (128, 128, 0, "scaled"), (256, 128, 1, "static")])
def test_theta_sums_time_dependent_v(gpu, nx, npp, tables, vkind):
    """The sums on the theta spectrum with a time-dependent V (the closing and
    opening factors of the fused column pass evaluated at different times) or
    a tabulated time-independent one, a non-polynomial K whose weights are
    not those of P and P^2, and the x-theta and lambda-p masks (which sit in the
    factors; no x-p mask, which would send the pass through x-p space).
    "field" is V0(X) + X F(t), split into a cached table of V0 and a per-step
    field factor; "scaled" is a time-dependent V that does not split and is
    evaluated at every half-step."""
    consts = {"E0": 0.05, "om": 0.3}
    masks = dict(abs_boundary_x_theta="pow(fabs(cos(0.5*M_PI*X_minus/X_amplitude)*cos(0.5*M_PI*X_plus/X_amplitude)), 0.01)",
                 abs_boundary_lambda_p="exp(-1e-4*Lambda*Lambda)")
    K = "0.5*P*P + 0.05*cos(P)"
    dK = "P - 0.05*sin(P)"
    V, dV = TD_V[vkind]
    g, o = make_pair(nx, 10., npp, 6., 0.02, V, K, dV, dK, consts=consts, masks=masks, phase_tables=tables)
    W0 = np.exp(-(o.X + 0.5) ** 2 - (o.P - 0.3) ** 2)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    Wg = g.propagate(200).get()
    o.propagate(200)
    assert rel_l2(Wg, o.W) <= 1e-12
    check_series(g, o)


@pytest.mark.parametrize("tables", [0, -1])
def test_fp32_form_error(gpu, tables):
    """FP32 W and FFTs (compensated FP32 twiddles), FP64 phase argument with its
    mod-2pi reduction, FP64 sums, against the FP64 oracle on the quartic 256^2 workload. Bar: 1e-4
    at 500 steps (the batched FP32 form's step count); 1000
    steps stay below 1.5e-4. For scale: plain float32 pocketfft drifts to 1.4e-4
    on the same 1000 steps."""
    g, o = make_pair(256, 8., 256, 8., 0.01, V_Q, K_Q, DV_Q, DK_Q, precision="fp32", phase_tables=tables)
    W0 = gaussian(o)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    errs = []
    for chunk in (500, 500):
        Wg = g.propagate(chunk).get()
        o.propagate(chunk)
        errs.append(rel_l2(Wg, o.W))
    print("quartic 256^2 fp32 (tables=%d) rel L2 vs fp64 oracle at 500/1000 steps:" % tables, errs)
    assert errs[0] <= 1e-4 and errs[1] <= 1.5e-4
    ox = np.array([s["X_average"] for s in o.series])
    assert np.max(np.abs(np.array(g.X_average) - ox)) <= 1e-4


def test_observable_getters(gpu):
    g, o = make_pair(128, 8., 128, 8., 0.01, V_Q, K_Q, DV_Q, DK_Q)
    W0 = gaussian(o)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    g.propagate(37)
    o.propagate(37)
    f = lambda X, P, t: X * P + 0.25 * X * X
    assert abs(g.get_average("X * P + 0.25 * X * X") - o.get_average(f)) <= 1e-12
    assert abs(g.get_purity() - o.get_purity()) <= 1e-12
    assert abs(g.get_sigma_x_sigma_p() - o.series[-1]["sigma_x_sigma_p"]) <= 1e-11


def test_set_wignerfunction_forms(gpu):
    """ndarray, C string, python function and float initial states agree."""
    from fastwigner_amd.wigner_moyal import WignerMoyalHIP1D
    kw = dict(X_gridDIM=64, X_amplitude=8., P_gridDIM=128, P_amplitude=6., dt=0.01, t=0., V="0.5*X*X", K="0.5*P*P")
    a = WignerMoyalHIP1D(**kw).set_wignerfunction("exp(-(X-1.)*(X-1.) - (P-0.5)*(P-0.5))").wignerfunction.get()
    sysb = WignerMoyalHIP1D(**kw)
    b = sysb.set_wignerfunction(lambda self, x, p: np.exp(-(x - 1.) ** 2 - (p - 0.5) ** 2)).wignerfunction.get()
    assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b))
    c = WignerMoyalHIP1D(**kw).set_wignerfunction(1.).wignerfunction.get()
    assert np.allclose(c, 1.0 / (64 * 128 * sysb.dXdP), rtol=1e-14)
    assert abs(b.sum() * sysb.dXdP - 1.0) <= 1e-13
