"""Strong-field Wigner propagation (the usage of strong_field_physics_cuda_1d.py:24-82
and strong_field_physics_fftw_1d.py:35-90 with the Wigner propagator): soft-core
V = -1/sqrt(x^2+1) + x E(t) with E(t) from `functions`,
E(t) = E0 sin^2(pi t/T) sin(omega t), E0 = 0.05, omega = 0.06, T = 4 cycles,
dt = 0.05, x in [-60, 60), p in [-6, 6), the x-p mask |cos(pi x / 2 X_amplitude)|^(1/8)
after every step, and the dipole acceleration -<V'(x) > (P_average_RHS, with
diff_V carrying E(t), strong_field_physics_cuda_1d.py:81) recorded every step.
The initial state is the field-free ground state from the Bloch propagator.
"""
import numpy as np
import pytest

from oracle import cexpr
from oracle import fw_oracle as fo

pytestmark = pytest.mark.gpu

OMEGA, E0 = 0.06, 0.05
T_PULSE = 4 * 2 * np.pi / OMEGA
FUNCTIONS = ("__device__ double E(double t)\n{{\n"
             "  return (t < T_pulse) ? E0 * pow(sin(M_PI * t / T_pulse), 2) * sin(omega * t) : 0.;\n}}\n")
V = "-1. / sqrt(X * X + 1.) + X * E(t)"
DV = "X / pow(X * X + 1., 1.5) + E(t)"
MASK = "pow(fabs(cos(0.5 * M_PI * X / X_amplitude)), 1. / 8.)"


def E_np(t):
    return np.where(t < T_PULSE, E0 * np.sin(np.pi * t / T_PULSE) ** 2 * np.sin(OMEGA * t), 0.)


def params(n, dt=0.05, **kw):
    d = dict(X_gridDIM=n, X_amplitude=60., P_gridDIM=n, P_amplitude=6., dt=dt, t=0., E0=E0, omega=OMEGA,
             T_pulse=T_PULSE, functions=FUNCTIONS, V=V, K="0.5 * P * P", diff_V=DV, diff_K="P",
             abs_boundary_x_p=MASK)
    d.update(kw)
    return d


def ground_state(n):
    """Field-free ground state by get_ground_state. The purity of the converged
    discrete state stays below 1 by about 1e-9 at 1024^2 (1e-7 at 256^2), so a
    tolerance of 1e-8 ends the loop there by purity, at 256^2 by the step size
    (each rejected step halves dbeta until it is below 1e-12)."""
    from fastwigner_amd.wigner_bloch import WignerBlochHIP1D
    b = WignerBlochHIP1D(dbeta=0.05, t_initial=0., **{k: v for k, v in params(n).items() if k != "abs_boundary_x_p"})
    return b.get_ground_state(abs_tol_purity=1e-8).get()


def test_strong_field_parity_256(gpu):
    """256^2, 400 steps in the pulse's rising edge with the field on, against the oracle."""
    from fastwigner_amd.wigner_moyal import WignerMoyalHIP1D
    n, steps, t0 = 256, 400, 0.25 * T_PULSE
    W0 = ground_state(n)
    g = WignerMoyalHIP1D(**params(n, t=t0))
    g.set_wignerfunction(W0)
    cx = dict(X_amplitude=60., dt=0.05)
    o = fo.MoyalOracle(n, 60., n, 6., 0.05, t=t0,
                       V=lambda X, t: -1. / np.sqrt(X * X + 1.) + X * E_np(t), K=lambda P, t: 0.5 * P * P,
                       diff_V=lambda X, t: X / (X * X + 1.) ** 1.5 + E_np(t), diff_K=lambda P, t: P,
                       mask_x_p=cexpr.make(MASK, cx, ("X", "P", "t")))
    o.set_wignerfunction(W0)
    W = g.propagate(steps).get()
    o.propagate(steps)
    err = np.linalg.norm(W - o.W) / np.linalg.norm(o.W)
    dip = np.array(g.P_average_RHS)
    dref = np.array([d["P_average_RHS"] for d in o.series])
    print("strong field 256^2 rel L2 %.2e, dipole max abs diff %.2e" % (err, np.max(np.abs(dip - dref))))
    assert err <= 1e-12
    assert np.max(np.abs(dip - dref)) <= 1e-11 * np.max(np.abs(dref))


def test_strong_field_full_pulse_1024_fp64_fp32(gpu, tmp_path):
    """The strong-field pulse at its full size: 1024^2, the whole 4-cycle pulse (8378 steps),
    FP64 and FP32 from the same ground state. Checked by what holds at any size:
    normalisation, the ground state barely moves before the pulse rises, a
    finite dipole series every step; and the FP32 form against FP64. The FP64
    run is saved the way the reference's driver saves it (settings, a frame per
    quarter of the pulse, the Ehrenfest series; strong_field_physics_cuda_1d.py:
    110-122,227-237,348-354), and the HHG spectrum of its dipole series
    (file_viwer.py:237-260) peaks at the laser frequency."""
    import time
    from fastwigner_amd.results import ResultsWriter, hhg_spectrum, load_results
    from fastwigner_amd.wigner_moyal import WignerMoyalHIP1D
    n = 1024
    steps = int(round(T_PULSE / 0.05))
    t0 = time.perf_counter()
    W0 = ground_state(n)
    print("strong field: ground state at %d^2 in %.1f s" % (n, time.perf_counter() - t0), flush=True)
    out = {}
    path = str(tmp_path / "strong_field.npz")
    for prec in ("fp64", "fp32"):
        t0 = time.perf_counter()
        g = WignerMoyalHIP1D(precision=prec, **params(n))
        g.set_wignerfunction(W0)
        if prec == "fp64":
            tw = 0.0
            with ResultsWriter(path, params(n), propagator=g) as rw:
                done = 0
                for q in range(4):
                    chunk = (steps * (q + 1)) // 4 - done
                    g.propagate(chunk)
                    done += chunk
                    t1 = time.perf_counter()
                    rw.add_propagator_frame(g)
                    tw += time.perf_counter() - t1
                rw.set_ehrenfest(g)
        else:
            g.propagate(steps)
        W = g.wignerfunction.get()
        print("strong field: %s pulse (%d steps) in %.1f s" % (prec, steps, time.perf_counter() - t0), flush=True)
        out[prec] = (W, np.array(g.P_average_RHS), np.array(g.X_average))
        assert abs(W.sum() * g.dX * g.dP - 1.0) <= (1e-12 if prec == "fp64" else 1e-6)
    W64, d64, x64 = out["fp64"]
    W32, d32, x32 = out["fp32"]
    assert len(d64) == steps and np.all(np.isfinite(d64))
    # the field is ~E0 (pi t/T)^2 omega t at the start: the ground state barely moves
    assert np.max(np.abs(x64[:20])) <= 1e-6
    werr = np.linalg.norm(W32 - W64) / np.linalg.norm(W64)
    derr = np.max(np.abs(d32 - d64)) / np.max(np.abs(d64))
    print("strong field: %d steps, FP32 vs FP64: W rel L2 %.2e, dipole max rel %.2e" % (steps, werr, derr))
    assert werr <= 1e-3
    assert derr <= 1e-3
    # the saved run reads back as written
    r = load_results(path)
    print("strong field: results file %.1f MiB, 4 frames of 8 MiB written in %.2f s"
          % (__import__("os").path.getsize(path) / 2 ** 20, tw))
    assert len(r["frames"]) == 4 and np.array_equal(r["frames"][-1]["wigner"], W64)
    assert abs(r["frames"][-1]["t"] - steps * 0.05) < 1e-9 and r["settings"]["X_gridDIM"] == n
    assert np.array_equal(r["ehrenfest"]["P_average_RHS"], d64)
    # HHG spectrum of the dipole acceleration: the driving frequency dominates
    om, spec = hhg_spectrum(d64, 0.05, OMEGA)
    pos = om > 0.5
    k = om[pos][np.argmax(spec[pos])]
    print("strong field: HHG spectrum peak at %.3f omega" % k)
    assert abs(k - 1.0) < 0.1


# velocity gauge (strong_field_physics_fftw_1d.py:35-80): Mathieu lattice, K(p, t) = (p + A(t))^2 / 2
VG_FUNCTIONS = ("__device__ double A(double t)\n{{\n"
                "  return -F / omega * sin(omega * t) * pow(sin(omega * t / 16.), 2);\n}}\n")


def vg_params(n, **kw):
    d = dict(X_gridDIM=n, X_amplitude=8., P_gridDIM=n, P_amplitude=10., dt=0.01, t=0., V0=0.37, F=0.04,
             omega=0.05698, kT=0.03, functions=VG_FUNCTIONS, V="-V0 * (1. + cos(2. * M_PI * X / X_amplitude))",
             diff_V="V0 * 2. * M_PI / X_amplitude * sin(2. * M_PI * X / X_amplitude)",
             K="0.5 * (P + A(t)) * (P + A(t))", diff_K="P + A(t)")
    d.update(kw)
    return d


def A_np(t, F=0.04, omega=0.05698):
    return -F / omega * np.sin(omega * t) * np.sin(omega * t / 16.) ** 2


def test_velocity_gauge_mathieu_parity(gpu):
    """The velocity-gauge usage of strong_field_physics_fftw_1d.py:35-80 at 256^2 (the
    script runs 512^2): Mathieu lattice V = -V0 (1 + cos(2 pi x / X_amplitude)), the
    time-dependent kinetic energy K(p, t) = (p + A(t))^2 / 2, so the K factors are
    evaluated at every half-step (no table), the FFTW time argument t_{n+1}
    (phase_time_shift = 0), and the initial state the FFTW-convention Gibbs state at
    kT = 0.03 (:155-159). 600 steps from t = 300, where the field is on, against the
    oracle; X_RHS = <p + A(t)> is the velocity-gauge current."""
    from fastwigner_amd.wigner_bloch import WignerBlochHIP1D
    from fastwigner_amd.wigner_moyal import WignerMoyalHIP1D
    n, t0, steps = 256, 300., 600
    b = WignerBlochHIP1D(bloch_convention="fftw", dbeta=0.01, t_initial=0.,
                         **{k: v for k, v in vg_params(n).items() if k != "t"})
    W0 = b.get_gibbs_state().get()
    assert b.gibbs_steps == 3333
    g = WignerMoyalHIP1D(phase_time_shift=0.0, **vg_params(n, t=t0))
    g.set_wignerfunction(W0)
    V0, L = 0.37, 8.
    o = fo.MoyalOracle(n, 8., n, 10., 0.01, t=t0, phase_time_shift=0.0,
                       V=lambda X, t: -V0 * (1. + np.cos(2. * np.pi * X / L)),
                       K=lambda P, t: 0.5 * (P + A_np(t)) ** 2,
                       diff_V=lambda X, t: V0 * 2. * np.pi / L * np.sin(2. * np.pi * X / L),
                       diff_K=lambda P, t: P + A_np(t))
    o.set_wignerfunction(W0)
    W = g.propagate(steps).get()
    o.propagate(steps)
    err = np.linalg.norm(W - o.W) / np.linalg.norm(o.W)
    cur = np.array(g.X_average_RHS)
    cref = np.array([d["X_average_RHS"] for d in o.series])
    print("velocity gauge 256^2: rel L2 %.2e, current max abs diff %.2e, max |A| %.3f"
          % (err, np.max(np.abs(cur - cref)), np.max(np.abs(A_np(t0 + 0.01 * np.arange(steps))))))
    assert err <= 1e-12
    assert np.max(np.abs(cur - cref)) <= 1e-11 * np.max(np.abs(cref))
