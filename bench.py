"""Benchmark: Wigner grid-point*steps/s of the plain FP64 Moyal step at 1024^2.

Workload (one "step" = one Strang step V/2-K-V/2 over the whole grid, plus the
per-step renormalisation and the logged observables <x>, <p>, <H>, purity):
quartic physics (x, p in [-8, 8), V = 0.5x^2 + 0.1x^4, K = p^2/2, dt = 0.01,
W0 = exp(-(x-1)^2 - (p-0.5)^2)/pi) on a 1024 x 1024 grid, FP64. The state is
resident in HBM before the timed region; nothing crosses PCIe inside it except
the per-chunk time/observable series (64 steps per chunk).

Prints ONE JSON line. Extra fields:
  roofline      dominant kernel's algorithmic HBM bytes per launch / its mean
                duration (HIP events on the propagator's stream) vs 8 TB/s
  cpu_baseline  the in-repo oracle (scipy.fft, pocketfft) on the host cores,
                timed on a bounded sample of the same workload
  parity        relative L2 of the GPU state against the oracle after the
                sample's steps (FP64 bar 1e-12)
  variants      FP32 form at 1024^2 and FP64 at 2048^2 with their own rates,
                roofline fractions and errors against the FP64 oracle; the
                strong-field pulse at 1024^2 (time-dependent V); the
                double-well von Neumann density-matrix step at 1024^2; the
                surrogate at 256^2; the batched random-collision workload
                (256 instances)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_PASS = {"fp64": 16, "fp32": 8}   # one read + one write of W per pass, per grid point

CFG = dict(X_amplitude=8., P_amplitude=8., dt=0.01, t=0.,
           V="0.5*X*X+0.1*X*X*X*X", K="0.5*P*P", diff_V="X+0.4*X*X*X", diff_K="P")


def w0(sys_):
    return np.exp(-(sys_.X - 1.) ** 2 - (sys_.P - 0.5) ** 2) / np.pi


def make(n, precision):
    from fastwigner_amd.wigner_moyal import WignerMoyalHIP1D
    s = WignerMoyalHIP1D(X_gridDIM=n, P_gridDIM=n, precision=precision, **CFG)
    s.set_wignerfunction(w0(s))
    return s


def kernel_times(s, reps=640):
    """Mean device time (us) per launch of the two-pass step kernels, fw_vv and
    fw_k_norm (fw_wigner_kernel_bench): the step time from HIP events around
    replays of the propagator's 64-step graph on its stream, split between the
    two kernels by their shares measured with an event around every kernel of
    one more chunk; the state is restored afterwards. {name: us}."""
    import ctypes
    from fastwigner_amd import _lib
    out = (ctypes.c_double * 2)()
    _lib.call("fw_wigner_kernel_bench", s._handle, float(s.t), int(reps), out)
    return {"fw_vv": out[0], "fw_k_norm": out[1]}


def measured_traffic(kernel, n, precision):
    """HBM-side bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (scaled by a copy kernel of
    known size), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "moyal_%s_%d_traffic.md" % (precision, n))
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return int(f[2])
    return None


def roofline(n, precision, ktimes, traffic=True):
    """Roofline entry of the dominant step kernel; `traffic` only for the quartic
    Moyal workloads the committed counter summaries cover."""
    name = max(ktimes, key=ktimes.get)
    bytes_per_launch = BYTES_PER_PASS[precision] * n * n
    achieved = bytes_per_launch / (ktimes[name] * 1e-6) / 1e9
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(name, n, precision) if traffic else None,
            "bytes_per_launch": bytes_per_launch, "passes_per_step": len(ktimes),
            "mean_us": {k: round(v, 3) for k, v in ktimes.items()}}


# strong-field Wigner propagation (strong_field_physics_cuda_1d.py:24-82 usage)
SF_OMEGA, SF_E0 = 0.06, 0.05
SF_T_PULSE = 4 * 2 * np.pi / SF_OMEGA
SF = dict(X_amplitude=60., P_amplitude=6., dt=0.05, t=0., E0=SF_E0, omega=SF_OMEGA, T_pulse=SF_T_PULSE,
            functions="__device__ double E(double t)\n{{\n  return (t < T_pulse) ? E0 * pow(sin(M_PI * t / T_pulse), 2)"
                      " * sin(omega * t) : 0.;\n}}\n",
            V="-1. / sqrt(X * X + 1.) + X * E(t)", K="0.5 * P * P", diff_V="X / pow(X * X + 1., 1.5) + E(t)",
            diff_K="P", abs_boundary_x_p="pow(fabs(cos(0.5 * M_PI * X / X_amplitude)), 1. / 8.)")


def strong_field_variant(n=1024, steps=2000, warmup=64):
    """Strong-field pulse at 1024^2, FP64: the field-free ground state (Bloch, get_ground_state),
    then `steps` steps inside the pulse with the time-dependent V = V0(X) + X E(t)
    (a cached table of V0's factors and the field phase exp(+i dt/2 Theta E) per
    half-step), the x-p mask after every step and the dipole acceleration
    P_average_RHS recorded every step."""
    from fastwigner_amd.wigner_bloch import WignerBlochHIP1D
    from fastwigner_amd.wigner_moyal import WignerMoyalHIP1D
    t0 = time.perf_counter()
    b = WignerBlochHIP1D(X_gridDIM=n, P_gridDIM=n, dbeta=0.05, t_initial=0.,
                         **{k: v for k, v in SF.items() if k != "abs_boundary_x_p"})
    W0 = b.get_ground_state(abs_tol_purity=1e-8).get()
    gs = time.perf_counter() - t0
    del b
    g = WignerMoyalHIP1D(X_gridDIM=n, P_gridDIM=n, **SF)
    g.set_wignerfunction(W0)
    rate, ms = timed_rate(g, n, steps, warmup)
    out = {"value": rate, "unit": "grid-point*steps/s", "ms_per_step": ms, "steps": steps,
           "t_window": [round(warmup * 0.05, 2), round((warmup + steps) * 0.05, 2)],
           "ground_state_s": round(gs, 2), "roofline": roofline(n, "fp64", kernel_times(g), traffic=False),
           "dipole_finite": bool(np.all(np.isfinite(g.P_average_RHS)))}
    print("[bench] strong field fp64 %dx%d: %.3e pt*steps/s, %.2f us/step" % (n, n, rate, ms * 1e3),
          file=sys.stderr, flush=True)
    return out


def rho_variant(n=1024, steps=2000, warmup=64):
    """Double-well density matrix (config 2): the thermal state at 1024^2
    (x in [-10, 10), kT = 0.2, RhoBloch from rho = I), kicked by exp(i (x - x')),
    then 2000 von Neumann steps (dt = 0.005) with the Ehrenfest series every
    step, in complex128 (three passes of 32 B per element: 96 B per element per
    step) and in complex64 with FP64 phases and sums (48 B), the latter's error
    against the FP64 run stated."""
    from fastwigner_amd.rho_bloch import RhoBlochHIP1D
    from fastwigner_amd.rho_vneumann import RhoVNeumannHIP1D
    V, K = "-0.5*X*X + X*X*X*X/20.", "0.5*P*P"
    b = RhoBlochHIP1D(X_gridDIM=n, X_amplitude=10., kT=0.2, dbeta=0.005, t=0., V=V, K=K)
    rho_th = b.get_gibbs_state().get()
    X = b.X
    del b
    # the thermal state with a momentum kick exp(i p0 (x - x')), so that the observables move
    rho0 = rho_th * np.exp(1j * 1.0 * (X[np.newaxis, :] - X[:, np.newaxis]))
    res = {}
    for prec, esz in (("fp64", 16), ("fp32", 8)):
        vn = RhoVNeumannHIP1D(X_gridDIM=n, X_amplitude=10., dt=0.005, t=0., V=V, K=K, diff_V="-X + X*X*X/5.",
                              diff_K="P", precision=prec)
        vn.set_rho(rho0)
        vn.propagate(warmup)
        t0 = time.perf_counter()
        vn.propagate(steps)
        el = time.perf_counter() - t0
        us = el * 1e6 / steps
        achieved = 6.0 * esz * n * n / (us * 1e-6) / 1e9
        res[prec] = dict(value=n * n * steps / el, ms=el * 1e3 / steps, achieved=achieved, esz=esz,
                         rho=vn.rho.get(), x=np.array(vn.X_average), p=np.array(vn.P_average),
                         h=np.array(vn.hamiltonian_average))
        print("[bench] von Neumann %s %dx%d: %.1f us/step" % (prec, n, n, us), file=sys.stderr, flush=True)
        del vn
    r64, r32 = res["fp64"], res["fp32"]
    out = {"value": r64["value"], "unit": "matrix-element*steps/s", "ms_per_step": r64["ms"], "steps": steps,
           "dtype": "c128", "state": "double-well thermal state (kT = 0.2) kicked by exp(i (x - x'))",
           "roofline": {"bound": "hbm", "kernel": "von Neumann step (three passes)",
                        "achieved": round(r64["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(r64["achieved"] / HBM_PEAK_GBS, 4), "traffic": None,
                        "bytes_per_step": 96 * n * n},
           "fp32": {"value": r32["value"], "ms_per_step": r32["ms"], "dtype": "c64 (FP64 phases and sums)",
                    "roofline_frac": round(r32["achieved"] / HBM_PEAK_GBS, 4),
                    "rel_l2_vs_fp64": float(np.linalg.norm(r32["rho"] - r64["rho"]) / np.linalg.norm(r64["rho"])),
                    "max_abs_err_x_average": float(np.max(np.abs(r32["x"] - r64["x"]))),
                    "max_abs_err_p_average": float(np.max(np.abs(r32["p"] - r64["p"]))),
                    "max_abs_err_hamiltonian_average": float(np.max(np.abs(r32["h"] - r64["h"]))),
                    "note": "<H> in FP32 carries the state's rounding noise at |P| up to pi/dX = 161, weighted by "
                            "P^2/2"}}
    return out


def timed_rate(s, n, steps, warmup):
    s.propagate(warmup)
    t0 = time.perf_counter()
    s.propagate(steps)          # returns after the stream is synchronised
    t1 = time.perf_counter()
    dt = t1 - t0
    return n * n * steps / dt, dt * 1e3 / steps


def host_info(workers):
    """What the CPU baseline ran on: logical CPUs the OS reports, the CPU model,
    the threads used and the library versions."""
    import platform
    import scipy
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "threads_used": workers,
            "threads_note": "the host share of a one-GPU box is 16 threads; scipy.fft takes them "
                            "(workers=), numpy's elementwise products run on one",
            "numpy": np.__version__, "scipy": scipy.__version__, "python": platform.python_version()}


def oracle_sample(n, max_seconds, max_steps, workers):
    """The oracle on the host cores: steps until max_seconds (bounded sample)."""
    import scipy  # noqa: F401
    from oracle import fw_oracle as fo
    o = fo.MoyalOracle(n, CFG["X_amplitude"], n, CFG["P_amplitude"], CFG["dt"],
                       V=lambda X, t: 0.5 * X * X + 0.1 * X ** 4, K=lambda P, t: 0.5 * P * P,
                       diff_V=lambda X, t: X + 0.4 * X ** 3, diff_K=lambda P, t: P, workers=workers)
    o.set_wignerfunction(w0(o))
    t0 = time.perf_counter()
    steps = 0
    while steps < max_steps and time.perf_counter() - t0 < max_seconds:
        o.propagate(1)
        steps += 1
    el = time.perf_counter() - t0
    return o, steps, el


def batched_variant(steps=500, o_workers=1, with_oracle=True):
    """Batched random-collision workload: 256 instances, grids from
    {128,256,512,1024}^2 (rng seed 4), quartic potentials, Gaussian initial
    states, gamma ~ U[0, 0.1], kT ~ U[0.1, 1] (Gibbs states from the batched
    Bloch run), FP32 with FP64 phases and sums, dt = 0.01, `steps` steps with
    one launch per pass over all instances, replayed from a hipGraph per 64
    steps. Rate in total
    grid-point*steps/s; error of the three smallest instances against the FP64
    oracle."""
    from fastwigner_amd.batch import WignerBatchHIP1D, gaussian_state, random_quartic_instances, QUARTIC_V, \
        QUARTIC_DV
    inst, _ = random_quartic_instances(256, seed=4)
    b = WignerBatchHIP1D(inst, V=QUARTIC_V, K="0.5*P*P", dt=0.01, param_names=("a2", "a3", "a4"),
                         diff_V=QUARTIC_DV, diff_K="P", precision="fp32")
    t0 = time.perf_counter()
    b.prepare_random_collision(dbeta=0.01)
    prep = time.perf_counter() - t0
    for i, d in enumerate(inst):
        b.set_wignerfunction(i, gaussian_state(d["X_gridDIM"], 8., d["x0"], d["p0"]))
    b.propagate(64, record=False)                     # warm-up (graph capture)
    for i, d in enumerate(inst):
        b.set_wignerfunction(i, gaussian_state(d["X_gridDIM"], 8., d["x0"], d["p0"]))
    b.t = 0.0
    t0 = time.perf_counter()
    b.propagate(steps)                                # observables of every instance, every step
    el = time.perf_counter() - t0
    rate = b.total_points * steps / el
    out = {"value": rate, "unit": "grid-point*steps/s", "ms_per_step": el * 1e3 / steps, "steps": steps,
           "instances": 256, "total_points": b.total_points, "gibbs_prep_s": round(prep, 2),
           "tasks_per_step": b.tasks_per_step, "persistent_grid": b.grid}
    # FP32 bytes a step must move: three passes read and write W (24 B per point) and the
    # dissipator reads W0s twice (8 B); the two-pass step reads and writes W twice (16 B), and a
    # dissipative instance also reads W0s and G = Vopen(W0s) once each (8 B)
    us = b.step_time_us(640)
    passes = b.passes_per_step
    pts = np.array([d["X_gridDIM"] ** 2 for d in inst], dtype=np.float64)
    dis = np.array([float(w) < 1.0 for w in b.w_coeff])
    bytes_step = int(np.sum(pts * ((16 if passes == 2 else 24) + 8 * dis)))
    out["passes_per_step"] = passes
    out["roofline"] = {"bound": "hbm", "kernel": "fw_batch_step (one step: %s)" % (
                           "one persistent launch" if b.mode == 0 else "%d launches, one per pass" % passes),
                       "achieved": round(bytes_step / (us * 1e-6) / 1e9, 1),
                       "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(bytes_step / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                       "bytes_per_step": bytes_step, "mean_us_per_step": round(us, 3)}
    print("[bench] batched fp32: %.3e pt*steps/s, %.1f us/step (device %.1f us)" % (rate, el * 1e6 / steps, us),
          file=sys.stderr, flush=True)
    if with_oracle:
        from oracle import fw_oracle as fo
        errs = []
        b2 = WignerBatchHIP1D(inst, V=QUARTIC_V, K="0.5*P*P", dt=0.01, param_names=("a2", "a3", "a4"),
                              diff_V=QUARTIC_DV, diff_K="P", precision="fp32")
        b2.prepare_random_collision(dbeta=0.01)
        for i, d in enumerate(inst):
            b2.set_wignerfunction(i, gaussian_state(d["X_gridDIM"], 8., d["x0"], d["p0"]))
        b2.propagate(steps, record=False)
        for i in sorted(range(256), key=lambda k: inst[k]["X_gridDIM"])[:3]:
            d = inst[i]
            a2, a3, a4 = d["params"]
            n = d["X_gridDIM"]
            o = fo.RandomCollisionOracle(n, 8., n, 8., 0.01, gamma=d["gamma"], kT=d["kT"], dbeta=0.01,
                                         V=lambda X, t: a2 * X * X + a3 * X ** 3 + a4 * X ** 4,
                                         K=lambda P, t: 0.5 * P * P, phase_time_shift=0.0, workers=o_workers)
            o.set_wignerfunction(gaussian_state(n, 8., d["x0"], d["p0"]))
            o.propagate(steps)
            W = b2.get_wignerfunction(i)
            errs.append(float(np.linalg.norm(W - o.W) / np.linalg.norm(o.W)))
        out["rel_l2_vs_fp64_oracle"] = {"instances": "three smallest", "max": max(errs), "values": errs}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run 10 launches of a plain elementwise copy (y = 1.0 * x) of 32 MiB, the "
                         "known-size reference that scales the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched FP32 random-collision workload (for profiling)")
    ap.add_argument("--rho-only", action="store_true",
                    help="only the double-well density-matrix workload, FP64 and FP32 (for profiling)")
    args = ap.parse_args()
    if args.rho_only:
        r = rho_variant(steps=args.steps)
        r.pop("rho", None)
        print(json.dumps(r), flush=True)
        return
    if args.batched_only:
        print(json.dumps(batched_variant(steps=args.steps, with_oracle=False)), flush=True)
        return
    if args.copy_probe:
        import torch
        x = torch.rand(2048 * 2048, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(10):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": x.numel() * 8, "copy_probe_bytes_written": y.numel() * 8,
                          "launches": 10}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: a single Wigner grid is kept on one GPU (sharding it would be "
              "transpose-bound over xGMI) and the benchmark measures one GPU; use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    n = args.n
    print("[bench] compiling and initialising %dx%d fp64" % (n, n), file=sys.stderr, flush=True)
    s = make(n, "fp64")
    # the roofline's kernel timing first (640 steps of graph replays, the state restored
    # afterwards): it also brings the GPU out of idle clocks before the warm-up and the
    # timed steps, which are then the usual W + K steps of the propagator
    kt = kernel_times(s)
    rl = roofline(n, "fp64", kt)
    rate, ms = timed_rate(s, n, args.steps, args.warmup)
    print("[bench] fp64 %dx%d: %.3e pt*steps/s, %.2f us/step" % (n, n, rate, ms * 1e3), file=sys.stderr, flush=True)

    result = {
        "metric": "Wigner grid-point*steps/s at 1024^2 (plain FP64 Moyal step)",
        "value": rate, "unit": "grid-point*steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Gaussian initial state W0 = exp(-(x-1)^2-(p-0.5)^2)/pi)",
        "config": {"workload": "plain FP64 Moyal split-operator step, %dx%d, x,p in [-8,8), "
                               "V=0.5x^2+0.1x^4, K=p^2/2, dt=0.01, renormalisation and <x>,<p>,<H>,purity "
                               "every step" % (n, n), "grid": [n, n], "parallelism": "single GPU"},
        "roofline": rl,
    }

    if not args.no_baseline:
        # the host's CPU share of a one-GPU box is 16 threads, whatever os.cpu_count() reports
        workers = min(os.cpu_count() or 1, 16)
        print("[bench] oracle sample on %d host threads" % workers, file=sys.stderr, flush=True)
        o, osteps, oel = oracle_sample(n, 15.0, 400, workers)
        result["cpu_baseline"] = {"value": n * n * osteps / oel, "unit": "grid-point*steps/s", "cores": workers,
                                  "kind": "port",
                                  "sample": "%d steps of the same %dx%d FP64 workload with the scipy.fft oracle "
                                            "(oracle/fw_oracle.py), %.1f s" % (osteps, n, n, oel),
                                  "host": host_info(workers)}
        _, osteps1, oel1 = oracle_sample(n, 6.0, 200, 1)
        result["cpu_baseline"]["single_thread"] = {
            "value": n * n * osteps1 / oel1, "unit": "grid-point*steps/s", "cores": 1,
            "sample": "%d steps, %.1f s" % (osteps1, oel1)}
        g = make(n, "fp64")
        g.propagate(osteps)
        err = float(np.linalg.norm(g.wignerfunction.get() - o.W) / np.linalg.norm(o.W))
        result["parity"] = {"rel_l2_vs_oracle": err, "steps": osteps, "tolerance": 1e-12}
        print("[bench] parity fp64 after %d steps: %.2e" % (osteps, err), file=sys.stderr, flush=True)
        del g
    else:
        o = None

    if not args.no_variants:
        var = {}
        s32 = make(n, "fp32")
        r32, ms32 = timed_rate(s32, n, args.steps, args.warmup)
        v = {"value": r32, "ms_per_step": ms32, "roofline": roofline(n, "fp32", kernel_times(s32))}
        if o is not None:
            g32 = make(n, "fp32")
            g32.propagate(osteps)
            v["rel_l2_vs_fp64_oracle"] = float(np.linalg.norm(g32.wignerfunction.get() - o.W) / np.linalg.norm(o.W))
            v["steps_compared"] = osteps
        var["fp32_%d" % n] = v
        print("[bench] fp32 %dx%d: %.3e pt*steps/s" % (n, n, r32), file=sys.stderr, flush=True)
        del s32
        s2 = make(2048, "fp64")
        r2, ms2 = timed_rate(s2, 2048, max(args.steps // 4, 100), 50)
        var["fp64_2048"] = {"value": r2, "ms_per_step": ms2, "roofline": roofline(2048, "fp64", kernel_times(s2))}
        print("[bench] fp64 2048^2: %.3e pt*steps/s" % r2, file=sys.stderr, flush=True)
        del s2
        var["strong_field_fp64_1024"] = strong_field_variant()
        var["rho_vneumann_c128_1024"] = rho_variant()
        print("[bench] surrogate: training on propagator trajectories", file=sys.stderr, flush=True)
        from fastwigner_amd.fno import surrogate_experiment
        var["surrogate_fno_256"] = surrogate_experiment(
            20000, held_out=512, epochs=3.0, base_steps=2, out_scale=1e-2, zero_mass=True,
            log=lambda m: print(m, file=sys.stderr, flush=True))
        var["batched_fp32_256_instances"] = batched_variant(o_workers=min(os.cpu_count() or 1, 16),
                                                      with_oracle=not args.no_baseline)
        result["variants"] = var

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
