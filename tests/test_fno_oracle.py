"""Pins the numpy FNO restatement (oracle/fno_oracle.py) before it checks the
HIP forward/backward kernels: central finite differences of the loss
(float64, no bf16 rounding), the matrix-form transforms against numpy's
rfft2 / irfft2, the spectral convolution with identity mode weights against a
direct DFT sum (no FFT), and the bf16 rounding rule."""
import numpy as np
import pytest

from oracle import fno_oracle as fo


def test_bf16_rounding_rule():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -7, -2.5, 3.14159265])
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3 2^-8 -> 1 + 2^-6; pi -> 3.140625
    assert np.array_equal(fo.bf16(x), [1.0, 1.0, 1.0 + 2 ** -6, 1.0 + 2 ** -7, -2.5, 3.140625])


def test_spectral_identity_is_lowpass_by_direct_dft():
    N, C, M = 8, 3, 2
    o = fo.FNOOracle(N=N, C=C, M=M, L=1, bf16=False)
    rng = np.random.default_rng(0)
    v = rng.standard_normal((1, C, N, N))           # channel images v[b][c][p][x]
    R = np.zeros((o.nm, C, C, 2))
    R[:, np.arange(C), np.arange(C), 0] = 1.0
    S = o.inverse(o.mix(o.modes(v), R))
    # direct: sum over the retained (kp, ky) of c_ky Re(Vhat e^{+i theta}) / N^2
    p = np.arange(N)[:, None]
    x = np.arange(N)[None, :]
    ref = np.zeros((C, N, N))
    for kp in list(range(M)) + list(range(N - M, N)):
        for ky in range(M):
            e = np.exp(-2j * np.pi * (kp * p + ky * x) / N)
            vh = np.einsum("px,cpx->c", e, v[0])
            ref += (1.0 if ky == 0 else 2.0) * np.real(vh[:, None, None] * np.conj(e)[None]) / N ** 2
    assert np.max(np.abs(S[0] - ref)) < 1e-12


def test_matrix_transforms_are_numpy_rfft2_and_irfft2():
    """Without rounding, the cosine/sine-matrix transforms are numpy's rfft2 on the
    retained modes and irfft2 of the truncated spectrum (N = 64, M = 8)."""
    N, C, M = 64, 3, 8
    o = fo.FNOOracle(N=N, C=C, M=M, L=1, bf16=False)
    rng = np.random.default_rng(7)
    v = rng.standard_normal((2, C, N, N))
    F = np.fft.rfft2(v, axes=(2, 3))[:, :, o.kp, :M]            # [b, c, 2M, M]
    Z = o.modes(v).transpose(0, 2, 1).reshape(2, C, 2 * M, M)
    assert np.max(np.abs(Z - F)) <= 1e-12 * np.max(np.abs(F))
    Y = rng.standard_normal((2, o.nm, C)) + 1j * rng.standard_normal((2, o.nm, C))
    full = np.zeros((2, C, N, N // 2 + 1), dtype=complex)
    full[:, :, o.kp, :M] = Y.transpose(0, 2, 1).reshape(2, C, 2 * M, M)
    ref = np.fft.irfft2(full, s=(N, N), axes=(2, 3))
    S = o.inverse(Y)
    assert np.max(np.abs(S - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_rounded_oracle_stays_close_to_the_exact_one():
    """The bf16 roundings (activations, matrices, intermediates, weights) move the
    network output by about the bf16 resolution, not more: a misplaced rounding of a
    large intermediate would show as a far larger change."""
    N, C, M, L = 64, 32, 8, 2
    ex = fo.FNOOracle(N=N, C=C, M=M, L=L, bf16=False)
    rd = fo.FNOOracle(N=N, C=C, M=M, L=L, bf16=True)
    rng = np.random.default_rng(5)
    P = fo.init_params(ex, seed=3, spectral_scale=3.0 / (C * C))
    P[ex.o_Q:ex.o_Q + C] = rng.standard_normal(C) * 0.05
    W = rng.random((2, N, N))
    feat = rng.random((2, 5))
    d_ex = ex.forward(P, W, feat) - W
    d_rd = rd.forward(P, W, feat) - W
    e = np.linalg.norm(d_rd - d_ex) / np.linalg.norm(d_ex)
    assert 1e-4 < e < 2e-2


@pytest.mark.parametrize("out_scale,zero_mass", [(1.0, False), (0.05, False), (0.05, True)])
def test_loss_gradient_by_finite_differences(out_scale, zero_mass):
    o = fo.FNOOracle(N=16, C=4, M=2, L=2, bf16=False, out_scale=out_scale, zero_mass=zero_mass)
    rng = np.random.default_rng(1)
    P = fo.init_params(o, seed=2, spectral_scale=0.3)
    P[o.o_Q:o.o_Q + o.C] = rng.standard_normal(o.C) * 0.5
    P[o.o_qb] = 0.01
    for l in range(o.L):
        P[o.o_b[l]:o.o_b[l] + o.C] = rng.standard_normal(o.C) * 0.1
    W = rng.random((2, 16, 16))
    feat = rng.random((2, 5))
    tgt = W + 0.1 * rng.standard_normal(W.shape)
    L0, g = o.loss_grad(P, W, feat, tgt)
    picks = np.concatenate([rng.choice(o.nparam, 60, replace=False), np.arange(o.o_Q, o.nparam),
                            np.arange(o.o_Pw, o.o_Pw + 8)])
    h = 1e-6
    for i in picks:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        num = (o.loss(o.forward(Pp, W, feat), tgt)[0] - o.loss(o.forward(Pm, W, feat), tgt)[0]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num)) + 1e-9, (i, num, g[i])


def test_planted_adjoint_mistake_is_caught():
    """The finite-difference check fails for a wrong adjoint (a dropped factor
    2 on the ky >= 1 modes of the inverse transform's adjoint)."""
    o = fo.FNOOracle(N=16, C=4, M=2, L=1, bf16=False)
    rng = np.random.default_rng(3)
    P = fo.init_params(o, seed=4, spectral_scale=0.3)
    P[o.o_Q:o.o_Q + o.C] = rng.standard_normal(o.C)
    W = rng.random((1, 16, 16))
    feat = rng.random((1, 5))
    tgt = W + 0.1 * rng.standard_normal(W.shape)
    i = o.o_R[0] + 1 * (o.C * o.C) * 2 + 1        # mode m = 1 (kp = 0, ky = 1): imaginary weight
    _, g = o.loss_grad(P, W, feat, tgt)
    Pp, Pm = P.copy(), P.copy()
    Pp[i] += 1e-6
    Pm[i] -= 1e-6
    num = (o.loss(o.forward(Pp, W, feat), tgt)[0] - o.loss(o.forward(Pm, W, feat), tgt)[0]) / 2e-6
    assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(num))
    o.cky = np.full(o.M, 1.0 / (o.N * o.N))         # the adjoint's weights without c_ky = 2
    _, g_bad = o.loss_grad(P, W, feat, tgt)
    assert abs(num - g_bad[i]) > 1e-3 * abs(num)


def test_zero_mass_output_keeps_the_mass_of_the_input():
    o = fo.FNOOracle(N=16, C=4, M=2, L=1, bf16=False, out_scale=0.3, zero_mass=True)
    rng = np.random.default_rng(3)
    P = fo.init_params(o, seed=1, spectral_scale=0.3)
    P[o.o_Q:o.o_Q + o.C] = rng.standard_normal(o.C)
    P[o.o_qb] = 0.7
    W = rng.random((3, 16, 16))
    out = o.forward(P, W, rng.random((3, 5)))
    assert np.allclose(out.sum(axis=(1, 2)), W.sum(axis=(1, 2)), rtol=0, atol=1e-12)
    assert np.max(np.abs(out - W)) > 1e-3          # and still corrects W


def test_energy_law_of_the_physical_referee():
    """The surrogate's referee checks <H> after a k-step jump against
    <H>_G + w^(2k) (<H>_0 - <H>_G) (fastwigner_amd.fno.energy_law). On the
    random-collision oracle (quartic V, gamma = 0.08) the law holds to the
    splitting error (below 1e-4 of the energy scale after 20 and 60 steps),
    well inside the referee's 1e-3; the slip w^k for w^(2k) misses by > 5e-3."""
    from oracle import fw_oracle as fo
    from fastwigner_amd.fno import energy_law, hamiltonian_grid, physical_referee, state_energy, state_moments
    a2, a3, a4, gamma, kT = 0.6, 0.05, 0.1, 0.08, 0.5
    n, amp, dt = 128, 8., 0.01
    V = lambda X, t: a2 * X * X + a3 * X ** 3 + a4 * X ** 4
    kw = dict(gamma=gamma, kT=kT, dbeta=0.01, V=V, K=lambda P, t: 0.5 * P * P, phase_time_shift=0.0)
    rc = fo.RandomCollisionOracle(n, amp, n, amp, dt, **kw)
    H = hamiltonian_grid(n, amp, a2, a3, a4)
    E = lambda W: (W * H).sum() / W.sum()
    W0 = np.exp(-(rc.X - 1.5) ** 2 - (rc.P + 0.5) ** 2)
    W0 /= W0.sum() * rc.dXdP
    e0, eg = E(W0), E(rc.gibbs)
    for k in (20, 60):
        r = fo.RandomCollisionOracle(n, amp, n, amp, dt, gibbs=rc.gibbs, **kw)
        r.set_wignerfunction(W0)
        r.propagate(k)
        assert abs(E(r.W) - energy_law(e0, eg, gamma, k, dt)) <= 1e-4 * max(abs(e0), 1.0)
        wrong = eg + np.exp(-gamma * dt / 2) ** k * (e0 - eg)
        assert abs(E(r.W) - wrong) > 5e-3 * max(abs(e0), 1.0)
        # the referee accepts the propagator's own answer and refuses a perturbed one
        feat = np.array([[a2, a3, a4, gamma, kT]])
        e_in = state_energy(W0[None], feat, amp)
        assert abs(e_in[0] - e0) <= 1e-12 * abs(e0)
        ok, drift, edrift, pur = physical_referee(state_moments(r.W[None]), e_in, feat, [eg], k, dt, amp)
        assert ok[0] and drift[0] < 1e-12 and edrift[0] < 1e-4 and pur[0] <= 1.0
        # the state moved by three momentum cells: norm and purity unchanged, the energy not
        bad = np.roll(r.W, 3, axis=0)
        okb, driftb, edriftb, purb = physical_referee(state_moments(bad[None]), e_in, feat, [eg], k, dt, amp)
        assert driftb[0] < 1e-12 and abs(purb[0] - pur[0]) < 1e-12
        assert edriftb[0] > 1e-3 and not okb[0]
