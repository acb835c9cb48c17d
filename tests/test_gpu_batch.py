"""Batched propagator (one persistent launch per step over many instances with
different grids, potentials, collision rates and temperatures) against the
single-grid oracle, instance by instance. FP64 bar 1e-12 relative L2; FP32
form 1e-4 against the FP64 oracle."""
import numpy as np
import pytest

from oracle import fw_oracle as fo

pytestmark = pytest.mark.gpu

V, DV = "a2*X*X + a3*X*X*X + a4*X*X*X*X", "2.*a2*X + 3.*a3*X*X + 4.*a4*X*X*X"


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def oracle_for(d, dt, gibbs_cache=None):
    a2, a3, a4 = d["params"]
    kw = dict(V=lambda X, t: a2 * X * X + a3 * X ** 3 + a4 * X ** 4, K=lambda P, t: 0.5 * P * P,
              diff_V=lambda X, t: 2 * a2 * X + 3 * a3 * X * X + 4 * a4 * X ** 3, diff_K=lambda P, t: P,
              phase_time_shift=0.0)
    n, A = d["X_gridDIM"], d["X_amplitude"]
    if d.get("gamma", 0) > 0:
        return fo.RandomCollisionOracle(n, A, n, A, dt, gamma=d["gamma"], kT=d["kT"], dbeta=0.01,
                                        bloch_convention="fftw", **kw)
    return fo.MoyalOracle(n, A, n, A, dt, **kw)


def small_batch():
    rng = np.random.default_rng(7)
    inst = []
    for k, n in enumerate((64, 128, 256, 128, 64, 256, 512)):
        inst.append(dict(X_gridDIM=n, X_amplitude=8., P_amplitude=8.,
                         params=(rng.uniform(0.2, 1), rng.uniform(-0.1, 0.1), rng.uniform(0, 0.2)),
                         gamma=(rng.uniform(0.02, 0.1) if k % 2 else 0.0), kT=rng.uniform(0.3, 1.0),
                         x0=rng.uniform(-2, 2), p0=rng.uniform(-2, 2)))
    return inst


def make_batch(inst, precision, mode=0, Vs=V):
    from fastwigner_amd.batch import WignerBatchHIP1D, gaussian_state
    b = WignerBatchHIP1D(inst, V=Vs, K="0.5*P*P", dt=0.01, param_names=("a2", "a3", "a4"), diff_V=DV, diff_K="P",
                         precision=precision, mode=mode)
    if any(d.get("gamma", 0) > 0 for d in inst):
        b.prepare_random_collision(dbeta=0.01)
    for i, d in enumerate(inst):
        b.set_wignerfunction(i, gaussian_state(d["X_gridDIM"], d["X_amplitude"], d["x0"], d["p0"]))
    return b


@pytest.mark.parametrize("mode,steps", [(0, 150), (1, 150), (1, 128), (1, 1), (1, 65)])
def test_batch_fp64_parity(gpu, mode, steps):
    """Seven instances (64^2 .. 512^2, Moyal and random collision with their own
    Gibbs states from the batched Bloch run), against the oracle per instance:
    150 steps (two graph chunks plus a partial one), and for the two-pass steps
    of mode 1 also runs ending exactly on a chunk, of one step and of one chunk
    plus one step."""
    from fastwigner_amd.batch import gaussian_state
    inst = small_batch()
    b = make_batch(inst, "fp64", mode)
    assert b.passes_per_step == (1 if mode == 0 else 2)
    b.propagate(steps)
    for i, d in enumerate(inst):
        o = oracle_for(d, 0.01)
        o.set_wignerfunction(gaussian_state(d["X_gridDIM"], d["X_amplitude"], d["x0"], d["p0"]))
        o.propagate(steps)
        W = b.get_wignerfunction(i)
        err = rel(W, o.W)
        obs = b.observables(i)
        xo = np.array([s["X_average"] for s in o.series])
        ho = np.array([s["hamiltonian_average"] for s in o.series])
        print("instance %d N=%d gamma=%.3f: rel L2 %.2e" % (i, d["X_gridDIM"], d.get("gamma", 0), err))
        assert err <= 1e-12, i
        assert np.max(np.abs(obs["X_average"] - xo)) <= 1e-11, i
        assert np.max(np.abs(obs["hamiltonian_average"] - ho)) <= 1e-11 * np.max(np.abs(ho)), i
        if d.get("gamma", 0) > 0:
            assert b.bloch_steps[i] == fo.dbeta_num_beta_steps(d["kT"], 0.01)[1]


def test_batch_fp32_form_and_modes_agree(gpu):
    """FP32 batch (FP64 phases and sums) against the FP64 oracle over 500 steps.
    The persistent launch (mode 0) is bit-identical to one launch per pass in
    three passes per step (mode 1 with a V that names t, which keeps the three
    passes); the two-pass steps (mode 1, V independent of t) meet the same
    oracle bar and agree with them to FP32 rounding."""
    from fastwigner_amd.batch import gaussian_state
    inst = small_batch()[:4]
    b0 = make_batch(inst, "fp32", 0)
    b1 = make_batch(inst, "fp32", 1, Vs=V + " + 0.*t")
    b2 = make_batch(inst, "fp32", 1)
    assert (b0.passes_per_step, b1.passes_per_step, b2.passes_per_step) == (1, 3, 2)
    for b in (b0, b1, b2):
        b.propagate(500)
    for i, d in enumerate(inst):
        W0, W1, W2 = b0.get_wignerfunction(i), b1.get_wignerfunction(i), b2.get_wignerfunction(i)
        assert np.array_equal(W0, W1), i
        assert np.array_equal(b0.series[:, i], b1.series[:, i]), i
        o = oracle_for(d, 0.01)
        o.set_wignerfunction(gaussian_state(d["X_gridDIM"], d["X_amplitude"], d["x0"], d["p0"]))
        o.propagate(500)
        err, err2, d02 = rel(W0, o.W), rel(W2, o.W), rel(W2, W0)
        print("fp32 instance %d N=%d: rel L2 vs fp64 oracle %.2e (two-pass %.2e; between them %.2e)"
              % (i, d["X_gridDIM"], err, err2, d02))
        assert err <= 1e-4 and err2 <= 1e-4, i
        # the two-pass steps take fewer FP32 transforms (sums on the theta
        # spectrum), so the two forms differ by about their own FP32 error
        assert d02 <= 1e-4, i
        assert np.max(np.abs(b2.series[:, i, 0] / b0.series[:, i, 0] - 1)) <= 1e-4, i


def test_batched_256_instances_full(gpu):
    """The batched random-collision workload at full size: 256 instances with
    grids from {128, 256, 512, 1024}^2 (rng seed 4), FP32, 500 steps, one launch
    per step. Size-independent checks on every instance (finite sums, purity
    <= 1 up to FP32 rounding), normalisation and the oracle on the three
    smallest instances."""
    from fastwigner_amd.batch import WignerBatchHIP1D, gaussian_state, random_quartic_instances
    inst, _ = random_quartic_instances(256, seed=4)
    b = WignerBatchHIP1D(inst, V=V, K="0.5*P*P", dt=0.01, param_names=("a2", "a3", "a4"), diff_V=DV,
                         diff_K="P", precision="fp32")
    b.prepare_random_collision(dbeta=0.01)
    # A drawn quartic with a4 near 0 and a3 != 0 is unbounded below on the
    # theta-extended grid (V(x -+ theta/2) is evaluated up to |x| ~ X_amplitude
    # + pi/(2 dP)); its Wigner-Bloch "Gibbs state" is then unphysical (purity >> 1)
    # in FP64 as well. Such instances are flagged, not used for the purity check.
    gibbs_purity = np.array([2 * np.pi * np.sum(b.get_wignerfunction(i) ** 2) * b.grids[i]["dX"] * b.grids[i]["dP"]
                             for i in range(256)])
    physical = gibbs_purity <= 1.0 + 1e-3
    print("256 instances: %d of 256 instances have an unphysical Gibbs state: %s"
          % ((~physical).sum(), np.flatnonzero(~physical).tolist()))
    assert physical.sum() >= 240
    for i, d in enumerate(inst):
        b.set_wignerfunction(i, gaussian_state(d["X_gridDIM"], 8., d["x0"], d["p0"]))
    b.propagate(500)
    assert b.series.shape == (500, 256, 9)
    over = 0.0
    for i in range(256):
        ob = b.observables(i)
        assert np.all(np.isfinite(b.series[:, i])), i
        if physical[i]:
            over = max(over, float(np.max(ob["purity"])) - 1.0)
    # purity <= 1 up to the FP32 form's rounding (its W error is ~5e-5 relative L2)
    print("256 instances: largest purity - 1 over the physical instances and all steps: %.2e" % over)
    assert over <= 1e-3
    small = sorted(np.flatnonzero(physical), key=lambda i: inst[i]["X_gridDIM"])[:3]
    for i in small:
        d = inst[i]
        W = b.get_wignerfunction(i)
        g = b.grids[i]
        assert abs(W.sum() * g["dX"] * g["dP"] - 1.0) <= 1e-5
        o = oracle_for(d, 0.01)
        o.set_wignerfunction(gaussian_state(d["X_gridDIM"], 8., d["x0"], d["p0"]))
        o.propagate(500)
        err = rel(W, o.W)
        print("256 instances, instance %d (N=%d, gamma=%.3f, kT=%.2f): fp32 rel L2 vs fp64 oracle %.2e"
              % (i, d["X_gridDIM"], d["gamma"], d["kT"], err))
        assert err <= 1e-4


def test_batch_two_pass_parametrised_nonpolynomial_k(gpu):
    """Two-pass steps with the sums on the theta spectrum, whose weights are
    built per instance: a K that depends on an instance parameter and is not a
    polynomial, with and without the dissipator, FP64 against the oracle (W and
    the logged sums, 100 steps)."""
    from fastwigner_amd.batch import WignerBatchHIP1D, gaussian_state
    inst = small_batch()[:4]
    Ks, dKs = "0.5*P*P + 0.2*a4*cos(P)", "P - 0.2*a4*sin(P)"
    b = WignerBatchHIP1D(inst, V=V, K=Ks, dt=0.01, param_names=("a2", "a3", "a4"), diff_V=DV, diff_K=dKs,
                         precision="fp64", mode=1)
    assert b.passes_per_step == 2
    b.prepare_random_collision(dbeta=0.01)
    for i, d in enumerate(inst):
        b.set_wignerfunction(i, gaussian_state(d["X_gridDIM"], d["X_amplitude"], d["x0"], d["p0"]))
    steps = 100
    b.propagate(steps)
    for i, d in enumerate(inst):
        a2, a3, a4 = d["params"]
        kw = dict(V=lambda X, t: a2 * X * X + a3 * X ** 3 + a4 * X ** 4,
                  K=lambda P, t: 0.5 * P * P + 0.2 * a4 * np.cos(P),
                  diff_V=lambda X, t: 2 * a2 * X + 3 * a3 * X * X + 4 * a4 * X ** 3,
                  diff_K=lambda P, t: P - 0.2 * a4 * np.sin(P), phase_time_shift=0.0)
        n, A = d["X_gridDIM"], d["X_amplitude"]
        if d.get("gamma", 0) > 0:
            o = fo.RandomCollisionOracle(n, A, n, A, 0.01, gamma=d["gamma"], kT=d["kT"], dbeta=0.01,
                                         bloch_convention="fftw", **kw)
        else:
            o = fo.MoyalOracle(n, A, n, A, 0.01, **kw)
        o.set_wignerfunction(gaussian_state(n, A, d["x0"], d["p0"]))
        o.propagate(steps)
        err = rel(b.get_wignerfunction(i), o.W)
        obs = b.observables(i)
        ho = np.array([s["hamiltonian_average"] for s in o.series])
        po = np.array([s["P_average"] for s in o.series])
        dko = np.array([s["X_average_RHS"] for s in o.series])
        print("instance %d N=%d gamma=%.3f: rel L2 %.2e" % (i, n, d.get("gamma", 0), err))
        assert err <= 1e-12, i
        assert np.max(np.abs(obs["hamiltonian_average"] - ho)) <= 1e-11 * np.max(np.abs(ho)), i
        assert np.max(np.abs(obs["P_average"] - po)) <= 1e-11, i
        if d.get("gamma", 0) == 0:   # the collision model's RHS adds its relaxation term
            assert np.max(np.abs(obs["X_average_RHS"] - dko)) <= 1e-11, i

