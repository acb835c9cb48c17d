"""Fourier-neural-operator surrogate of k split-operator steps, trained on the
GPU by the library's own forward/backward kernels (csrc/fw_fno.cpp) on
trajectories of the batched propagator, with the plain propagator as the
referee and fallback.

The reference has no learned component: this is new work on top of the
split-operator step it defines (wigner_moyal_cuda_1d.py:496-540 and the
random-collision step, wigner_random_collision_model_fftw_1d.py:56-76).

Network: lift 8 -> C, L Fourier layers (spectral convolution on M x M retained
modes plus pointwise C x C mixing, GELU), projection C -> 1 added to the input
W (the untrained network, with a zero projection, is the identity map).
Each layer is three fused kernels on BF16 MFMA with FP32 accumulation: the
truncated forward DFT of each channel image, the per-mode channel mixing, and
the inverse fused with the pointwise mixing, bias and GELU (the projection in
the last layer). Activations are stored as bf16, channel-first.

Base method and learned correction: with base_steps = b > 0 the input of the
network is not W(t) but the propagator's own coarse prediction, b split-operator
steps of length k dt / b from W(t) (the same batched propagator, FP32), and the
network learns the correction that takes it to W(t + k dt) of the k fine steps.
The untrained network is then exactly the coarse method, so what the training
gains is measured against the method without the network. base_steps = 0 is the
plain jump from W(t).
"""
import ctypes

import numpy as np

from . import _lib
from .batch import QUARTIC_DV, QUARTIC_V, WignerBatchHIP1D, gaussian_state, random_quartic_instances


class FNOSurrogate:
    def __init__(self, grid_dim=256, width=32, modes=16, layers=4, max_batch=8, k_steps=20, dt=0.01, amplitude=8.,
                 seed=0, device=0, base_steps=0, out_scale=1.0, zero_mass=False):
        self.N, self.C, self.M, self.L = int(grid_dim), int(width), int(modes), int(layers)
        self.max_batch, self.k, self.dt, self.amplitude = int(max_batch), int(k_steps), float(dt), float(amplitude)
        self.base_steps = int(base_steps)
        if self.base_steps < 0 or (self.base_steps and self.k % self.base_steps):
            raise ValueError("base_steps must be 0 or divide k_steps")
        d = _lib.FnoDesc()
        d.grid_dim, d.width, d.modes, d.layers, d.max_batch, d.device = self.N, self.C, self.M, self.L, \
            self.max_batch, int(device)
        d.x_amplitude = d.p_amplitude = self.amplitude
        h = ctypes.c_void_p()
        _lib.call("fw_fno_create", ctypes.byref(d), ctypes.byref(h))
        self._handle = h
        n = ctypes.c_longlong()
        _lib.call("fw_fno_param_count", self._handle, ctypes.byref(n))
        self.nparam = n.value
        self.out_scale, self.zero_mass = float(out_scale), bool(zero_mass)
        _lib.call("fw_fno_set_output", self._handle, ctypes.c_float(self.out_scale), 1 if self.zero_mass else 0)
        self.set_params(self.initial_params(seed))
        self.npairs = 0
        self.history = []

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h:
            try:
                _lib.load().fw_fno_destroy(h)
            except Exception:
                pass
            self._handle = None

    # -- parameters ------------------------------------------------------------------
    def layout(self):
        C, nm = self.C, 2 * self.M * self.M
        o, out = 0, {}
        out["Pw"] = (o, 8 * C); o += 8 * C
        out["pb"] = (o, C); o += C
        for l in range(self.L):
            out["R%d" % l] = (o, nm * C * C * 2); o += nm * C * C * 2
            out["Wp%d" % l] = (o, C * C); o += C * C
            out["b%d" % l] = (o, C); o += C
        out["Q"] = (o, C); o += C
        out["qb"] = (o, 1); o += 1
        return out

    def initial_params(self, seed):
        """Lift and pointwise weights ~ N(0, 1/fan_in); spectral weights ~ U(-1/C^2, 1/C^2);
        biases and projection zero."""
        rng = np.random.default_rng(seed)
        P = np.zeros(self.nparam, dtype=np.float32)
        lay = self.layout()
        o, n = lay["Pw"]
        P[o:o + n] = rng.standard_normal(n) / np.sqrt(8.0)
        s = 1.0 / (self.C * self.C)
        for l in range(self.L):
            o, n = lay["R%d" % l]
            P[o:o + n] = rng.uniform(-s, s, n)
            o, n = lay["Wp%d" % l]
            P[o:o + n] = rng.standard_normal(n) / np.sqrt(self.C)
        return P

    def set_params(self, P):
        P = np.ascontiguousarray(P, dtype=np.float32)
        assert P.size == self.nparam
        _lib.call("fw_fno_set_params", self._handle, _lib.fptr(P))

    def get_params(self):
        P = np.empty(self.nparam, dtype=np.float32)
        _lib.call("fw_fno_get_params", self._handle, _lib.fptr(P))
        return P

    # -- inference and gradients (host arrays) ---------------------------------------------
    def predict(self, W, feat, moments=False):
        """W(t + k dt) for a batch: W [B, N, N], feat [B, 5] = (a2, a3, a4, gamma, kT).
        moments=True also returns the predictions' moments computed on the device
        (fw_fno_moments: x marginal, p marginal, row sums of W^2; [B, 3N] FP64)."""
        W = np.ascontiguousarray(W, dtype=np.float32)
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        out = np.empty_like(W)
        mom = np.empty((W.shape[0], 3 * self.N)) if moments else None
        for s in range(0, W.shape[0], self.max_batch):
            e = min(s + self.max_batch, W.shape[0])
            o = np.empty((e - s,) + W.shape[1:], dtype=np.float32)
            _lib.call("fw_fno_forward", self._handle, _lib.fptr(np.ascontiguousarray(W[s:e])),
                      _lib.fptr(np.ascontiguousarray(feat[s:e])), e - s, _lib.fptr(o))
            out[s:e] = o
            if moments:
                m = np.empty((e - s, 3 * self.N))
                _lib.call("fw_fno_moments", self._handle, e - s, _lib.dptr(m))
                mom[s:e] = m
        return (out, mom) if moments else out

    def moments_time_us(self, batch=None, reps=20):
        """Wall time of fw_fno_moments for one batch (after a forward), copy to the host included."""
        import time
        batch = batch or self.max_batch
        m = np.empty((batch, 3 * self.N))
        _lib.call("fw_fno_moments", self._handle, batch, _lib.dptr(m))
        t0 = time.perf_counter()
        for _ in range(reps):
            _lib.call("fw_fno_moments", self._handle, batch, _lib.dptr(m))
        return (time.perf_counter() - t0) * 1e6 / reps

    def loss_grad(self, W, feat, target, want_grad=True):
        W = np.ascontiguousarray(W, dtype=np.float32)
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        loss = ctypes.c_double()
        g = np.empty(self.nparam, dtype=np.float32) if want_grad else None
        _lib.call("fw_fno_loss_grad", self._handle, _lib.fptr(W), _lib.fptr(feat), _lib.fptr(target), W.shape[0],
                  ctypes.byref(loss), _lib.fptr(g) if want_grad else None)
        return loss.value, g

    # -- the base method ------------------------------------------------------------------
    def coarse_batch(self, instances, precision="fp32", prepare_dbeta=0.01):
        """The batched random-collision propagator with the coarse step k dt / b
        (Gibbs sources prepared), or None for base_steps = 0."""
        if not self.base_steps:
            return None
        c = WignerBatchHIP1D(instances, V=QUARTIC_V, K="0.5*P*P", dt=self.k * self.dt / self.base_steps,
                             param_names=("a2", "a3", "a4"), diff_V=QUARTIC_DV, diff_K="P", precision=precision)
        c.prepare_random_collision(dbeta=prepare_dbeta)
        return c

    # -- training data from the batched propagator --------------------------------------------
    def build_dataset(self, n_instances, pairs_per_instance, seed=4, prepare_dbeta=0.01):
        """Trajectories of the batched random-collision propagator (FP32) on N x N
        grids: n_instances seeded quartic systems (the same generator as the
        batched workload, grid fixed to N), each contributing the pairs
        (W(j k dt), W((j+1) k dt)) for j < pairs_per_instance, the input being
        the coarse base prediction from W(j k dt) when base_steps > 0. Instances
        whose Wigner-Bloch Gibbs state is unphysical (purity > 1; quartic
        unbounded below on the theta-extended grid) are left out. Returns the
        kept instances."""
        inst, _ = random_quartic_instances(n_instances, seed=seed, sizes=(self.N,), amplitude=self.amplitude)
        b = WignerBatchHIP1D(inst, V=QUARTIC_V, K="0.5*P*P", dt=self.dt, param_names=("a2", "a3", "a4"),
                             diff_V=QUARTIC_DV, diff_K="P", precision="fp32")
        b.prepare_random_collision(dbeta=prepare_dbeta)
        c = self.coarse_batch(inst, prepare_dbeta=prepare_dbeta)
        dXdP = (2 * self.amplitude / self.N) ** 2
        keep = []
        for i in range(len(inst)):
            Wg = b.get_wignerfunction(i)
            if 2 * np.pi * np.sum(Wg * Wg) * dXdP <= 1.0 + 1e-3:
                keep.append(i)
        for i, d in enumerate(inst):
            b.set_wignerfunction(i, gaussian_state(self.N, self.amplitude, d["x0"], d["p0"]))
        total = len(keep) * pairs_per_instance
        _lib.call("fw_fno_dataset_reserve", self._handle, total)
        def state_ptrs(batch):
            out = []
            for i in keep:
                p = ctypes.c_void_p()
                _lib.call("fw_batch_state_ptr", batch._handle, i, ctypes.byref(p), None)
                out.append(p)
            return out
        ptrs = state_ptrs(b)
        inptrs = ptrs if c is None else state_ptrs(c)
        feats = np.array([[inst[i]["params"][0], inst[i]["params"][1], inst[i]["params"][2], inst[i]["gamma"],
                           inst[i]["kT"]] for i in keep], dtype=np.float32)
        for j in range(pairs_per_instance):
            if c is not None:
                c.copy_states_from(b)
                c.propagate(self.base_steps, record=False)
            for q, i in enumerate(keep):
                _lib.call("fw_fno_dataset_put", self._handle, j * len(keep) + q, inptrs[q], None,
                          _lib.fptr(np.ascontiguousarray(feats[q])))
            b.propagate(self.k, record=False)
            for q, i in enumerate(keep):
                _lib.call("fw_fno_dataset_put", self._handle, j * len(keep) + q, None, ptrs[q],
                          _lib.fptr(np.ascontiguousarray(feats[q])))
        self.npairs = total
        self.dataset_instances = [inst[i] for i in keep]
        self.dataset_feats = feats
        del b, c
        return self.dataset_instances

    # -- training ----------------------------------------------------------------------
    def train(self, steps, lr=1e-3, batch=None, seed=0, pairs=None, lr_final=None, log_every=0):
        """Adam steps on random minibatches of the training pairs (indices `pairs`,
        default all); cosine decay from lr to lr_final."""
        batch = batch or self.max_batch
        rng = np.random.default_rng(seed)
        pool = np.arange(self.npairs) if pairs is None else np.asarray(pairs)
        lr_final = lr * 0.05 if lr_final is None else lr_final
        loss = ctypes.c_double()
        for s in range(steps):
            idx = np.ascontiguousarray(rng.choice(pool, size=batch, replace=False).astype(np.int32))
            lr_s = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * s / max(steps - 1, 1)))
            want = log_every and (s % log_every == 0 or s == steps - 1)
            _lib.call("fw_fno_train_step", self._handle, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), batch,
                      ctypes.c_float(lr_s), ctypes.byref(loss) if want else None)
            if want:
                self.history.append((s, loss.value))
        return self.history

    def eval_pairs(self, pairs):
        pairs = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32))
        out = np.empty(len(pairs))
        _lib.call("fw_fno_eval_pairs", self._handle, pairs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(pairs),
                  _lib.dptr(out))
        return out

    def infer_time_us(self, batch=None, reps=20):
        us = ctypes.c_double()
        _lib.call("fw_fno_infer_time", self._handle, batch or self.max_batch, reps, ctypes.byref(us))
        return us.value


def checked_jump(surrogate, W0, feat, ref_states, dXdP, l2_tol=1e-3, norm_tol=1e-4):
    """One k-step jump of each sample with the surrogate, checked against the plain
    propagator's result: accepted when the relative L2 error is <= l2_tol and the
    norm drift |sum(W) dXdP - 1| <= norm_tol, otherwise replaced by the referee.
    Returns (states, accepted mask, relative L2 errors, norm drifts). The norm
    drift is that of the raw network output; the accepted states are then
    renormalised like every propagator step (W /= sum(W) dXdP)."""
    pred = surrogate.predict(W0, feat).astype(np.float64)
    err = np.linalg.norm((pred - ref_states).reshape(len(pred), -1), axis=1) / \
        np.linalg.norm(ref_states.reshape(len(pred), -1), axis=1)
    drift = np.abs(pred.reshape(len(pred), -1).sum(axis=1) * dXdP - 1.0)
    ok = (err <= l2_tol) & (drift <= norm_tol)
    pred = pred / (pred.reshape(len(pred), -1).sum(axis=1)[:, None, None] * dXdP)
    out = np.where(ok[:, None, None], pred, ref_states)
    return out, ok, err, drift


def forward_flops(batch, n=256, width=32, modes=16, layers=4):
    """FLOPs (2 M N K per product) of one forward over `batch` inputs, stage by
    stage as csrc/fw_fno.cpp runs them, all on BF16 MFMA: per layer the forward
    DFT's x stage (2M outputs per row of each channel image) and p stage, the
    per-mode channel mixing, the inverse's p stage (2 M C outputs per row from 4M
    inputs) and x stage, and the pointwise C x C mixing."""
    C, M, N, B = width, modes, n, batch
    nm = 2 * M * M
    stage_a = 2 * (2 * M) * N * N * (B * C)
    stage_b = 2 * (4 * M) * (2 * M) * N * (B * C)
    mixing = 2 * B * (2 * C) * (2 * C) * nm
    inv_p = 2 * N * (2 * M * C) * (4 * M) * B
    inv_x = 2 * N * N * (2 * M) * C * B
    pointwise = 2 * B * N * N * C * C
    return layers * (stage_a + stage_b + mixing + inv_p + inv_x + pointwise)


def forward_bytes(batch, n=256, width=32, modes=16, layers=4, zero_mass=True):
    """Algorithmic HBM bytes of one forward (each tensor read or written once per
    kernel): layer 0's DFT reads W (FP32) and writes v0 (bf16), the others read
    v (bf16); the DFT writes the spectrum Z (FP32), the mixing reads Z and the
    bf16 mode weights and writes the inverse's operand BT (bf16); the inverse
    reads BT and v and writes the next v (bf16), the last layer W and the
    prediction (FP32) instead; the zero-mass correction reads and writes the
    prediction once more."""
    C, M, N, B = width, modes, n, batch
    nm = 2 * M * M
    act, img = B * C * N * N * 2, B * N * N * 4
    z, bt, wm = B * nm * 2 * C * 4, B * (2 * M * C) * (4 * M) * 2, nm * (2 * C) * (2 * C) * 2
    total = 0
    for l in range(layers):
        total += (img + act) if l == 0 else act
        total += z + z + wm + bt
        total += bt + act + (act if l < layers - 1 else 2 * img)
    return total + (2 * img if zero_mass else 0)


def hamiltonian_grid(n, amplitude, a2, a3, a4):
    """H(x, p) = p^2/2 + a2 x^2 + a3 x^3 + a4 x^4 on the n x n grid (rows p, columns x)."""
    d = 2. * amplitude / n
    x = d * (np.arange(n) - n / 2)
    return (a2 * x ** 2 + a3 * x ** 3 + a4 * x ** 4)[np.newaxis, :] + (0.5 * x ** 2)[:, np.newaxis]


def energy_law(e_in, e_gibbs, gamma, k, dt):
    """<H> after k random-collision steps of the split operator: each step is
    dissipator, Moyal step, dissipator, the dissipator W <- w W + (1 - w) W_G
    (w = exp(-gamma dt/2), wigner_random_collision_model_fftw_1d.py:32,63-74)
    moves <H> to w <H> + (1 - w) <H>_G and the Moyal step keeps <H> (to the
    splitting error), so <H>_k = <H>_G + w^(2k) (<H>_0 - <H>_G)."""
    w = np.exp(-np.asarray(gamma) * dt / 2.)
    return e_gibbs + w ** (2 * k) * (e_in - e_gibbs)


def state_moments(W):
    """The moments fw_fno_moments computes on the device, here from host arrays:
    [B, 3N] = [x marginal | p marginal | row sums of W^2]."""
    W = np.asarray(W, dtype=np.float64)
    return np.concatenate([W.sum(axis=1), W.sum(axis=2), (W * W).sum(axis=2)], axis=1)


def physical_referee(mom, e_in, feat, e_gibbs, k, dt, amplitude, norm_tol=1e-4, energy_tol=1e-3):
    """Referee of k-step jumps that needs no reference propagation, from the
    predictions' moments (state_moments / fw_fno_moments, [B, 3N]): a
    prediction is accepted when its norm drift |sum(W) dXdP - 1| <= norm_tol,
    its energy drift |<H>_pred - <H>_law| / max(|<H>_0|, 1) <= energy_tol
    against the relaxation law of the random-collision step (energy_law, from
    e_in = <H> of the states the jumps start from), and its purity
    2 pi sum(W^2) dXdP / norm^2 <= 1. The energies come from the marginals
    (H = K(p) + V(x)). Returns (ok, norm drift, energy drift, purity)."""
    mom = np.asarray(mom, dtype=np.float64)
    feat = np.asarray(feat, dtype=np.float64)
    n = mom.shape[1] // 3
    d = 2. * amplitude / n
    dXdP = d * d
    x = d * (np.arange(n) - n / 2)
    V = feat[:, 0:1] * x ** 2 + feat[:, 1:2] * x ** 3 + feat[:, 2:3] * x ** 4       # (B, N), along x
    K = 0.5 * x ** 2                                                                  # along p
    mx, mp, sq = mom[:, :n], mom[:, n:2 * n], mom[:, 2 * n:]
    tot = mx.sum(axis=1)
    e_pred = ((mx * V).sum(axis=1) + mp @ K) / tot
    s = tot * dXdP
    drift = np.abs(s - 1.0)
    e_in = np.asarray(e_in, dtype=np.float64)
    edrift = np.abs(e_pred - energy_law(e_in, np.asarray(e_gibbs), feat[:, 3], k, dt)) / np.maximum(np.abs(e_in), 1.0)
    pur = 2 * np.pi * sq.sum(axis=1) * dXdP / (s * s)
    ok = (drift <= norm_tol) & (edrift <= energy_tol) & (pur <= 1.0)
    return ok, drift, edrift, pur


def state_energy(W, feat, amplitude):
    """<H> of host states [B, N, N] for the potentials in feat (H = p^2/2 + V)."""
    mom = state_moments(W)
    feat = np.asarray(feat, dtype=np.float64)
    n = mom.shape[1] // 3
    d = 2. * amplitude / n
    x = d * (np.arange(n) - n / 2)
    V = feat[:, 0:1] * x ** 2 + feat[:, 1:2] * x ** 3 + feat[:, 2:3] * x ** 4
    mx, mp = mom[:, :n], mom[:, n:2 * n]
    return ((mx * V).sum(axis=1) + mp @ (0.5 * x ** 2)) / mx.sum(axis=1)


def surrogate_experiment(n_pairs=20000, pairs_per_instance=20, held_out=512, epochs=2.0, batch=8, infer_batch=32, lr=2e-3,
                         seed=4, held_seed=5, base_steps=2, out_scale=1e-2, zero_mass=True, log=None):
    """The learned-surrogate workload: a 4-layer, 32-channel, 16 x 16-mode FNO on
    256^2 Wigner functions jumping k = 20 steps of dt = 0.01, trained on about
    n_pairs trajectory pairs of the batched random-collision propagator (random
    quartic potentials, Gaussian initial states, seeded), then evaluated on
    `held_out` new seeds against the FP64 split-operator referee (the batched
    propagator in FP64), with fallback where the relative L2 error exceeds 1e-3
    or the norm drift 1e-4. With base_steps = b > 0 the network corrects the
    coarse prediction of b propagator steps of k dt / b (FP32), whose own error
    is reported as the method without the network, and its output is scaled by
    out_scale (the correction is small); with zero_mass the correction carries
    no mass, so the prediction keeps the base method's norm. Training takes
    minibatches of `batch`; the rates are for inference over `infer_batch`
    samples per forward (per-sample results do not depend on the batch). Returns a
    dict of measurements."""
    import time
    say = log or (lambda *a: None)
    s = FNOSurrogate(grid_dim=256, width=32, modes=16, layers=4, max_batch=max(batch, infer_batch), k_steps=20, dt=0.01, seed=0,
                     base_steps=base_steps, out_scale=out_scale if base_steps else 1.0, zero_mass=zero_mass)
    n_inst = int(np.ceil(n_pairs / pairs_per_instance / 0.97))
    t0 = time.perf_counter()
    kept = s.build_dataset(n_inst, pairs_per_instance, seed=seed)
    t_data = time.perf_counter() - t0
    say("[fno] dataset: %d pairs from %d instances in %.1f s" % (s.npairs, len(kept), t_data))
    steps = int(epochs * s.npairs / batch)
    t0 = time.perf_counter()
    s.train(steps, lr=lr, batch=batch, log_every=max(steps // 10, 1))
    t_train = time.perf_counter() - t0
    say("[fno] %d Adam steps in %.1f s, loss %s" % (steps, t_train, [round(l, 5) for _, l in s.history]))
    # held-out seeds and the FP64 referee
    inst, _ = random_quartic_instances(held_out, seed=held_seed, sizes=(256,), amplitude=8.)
    referee = WignerBatchHIP1D(inst, V=QUARTIC_V, K="0.5*P*P", dt=0.01, param_names=("a2", "a3", "a4"),
                             diff_V=QUARTIC_DV, diff_K="P", precision="fp64")
    referee.prepare_random_collision(dbeta=0.01)
    dXdP = (16. / 256) ** 2
    gibbs = [referee.get_wignerfunction(i) for i in range(held_out)]
    phys = [i for i in range(held_out) if 2 * np.pi * np.sum(gibbs[i] ** 2) * dXdP <= 1.0 + 1e-3]
    # <H> of each Gibbs state (the dissipator's target), for the referee's energy law
    e_gibbs = np.array([(gibbs[i] * hamiltonian_grid(256, 8., *inst[i]["params"])).sum() * dXdP /
                        (gibbs[i].sum() * dXdP) for i in phys])
    del gibbs
    W0 = np.stack([gaussian_state(256, 8., d["x0"], d["p0"]) for d in inst])
    for i in range(held_out):
        referee.set_wignerfunction(i, W0[i])
    t0 = time.perf_counter()
    referee.propagate(20, record=False)
    t_referee = time.perf_counter() - t0
    WJ = np.stack([referee.get_wignerfunction(i) for i in phys])
    feat = np.array([[inst[i]["params"][0], inst[i]["params"][1], inst[i]["params"][2], inst[i]["gamma"],
                      inst[i]["kT"]] for i in phys])
    W0p = W0[phys] / (W0[phys].sum(axis=(1, 2), keepdims=True) * dXdP)
    base_us = 0.0
    if base_steps:
        c = s.coarse_batch(inst)
        for i in range(held_out):
            c.set_wignerfunction(i, W0[i])
        c.propagate(base_steps, record=False)
        Xin = np.stack([c.get_wignerfunction(i) for i in phys])
        # device time of the base for one network batch: its share of the coarse batch's steps
        base_us = c.step_time_us(64) * base_steps * infer_batch / held_out
        del c
    else:
        Xin = W0p
    _, ok, err, drift = checked_jump(s, Xin, feat, WJ, dXdP)
    rel = lambda A: np.linalg.norm((A - WJ).reshape(len(phys), -1), axis=1) / np.linalg.norm(
        WJ.reshape(len(phys), -1), axis=1)
    ident = rel(W0p)
    base_err = rel(Xin)
    # the referee without the FP64 answer, on the same predictions, from their moments computed on
    # the device; its cost per network batch: the moments (device + copy) and the host's O(N) checks
    e_in = state_energy(W0p, feat, 8.)
    _, mom = s.predict(Xin, feat, moments=True)
    t0 = time.perf_counter()
    pok, pdrift, pedrift, ppur = physical_referee(mom, e_in, feat, e_gibbs, 20, 0.01, 8.)
    phys_us = (time.perf_counter() - t0) * 1e6 * infer_batch / len(phys) + s.moments_time_us(infer_batch)
    # the same law on the FP64 answers: the splitting's own energy drift, the floor of the criterion
    _, _, ref_edrift, _ = physical_referee(state_moments(WJ), e_in, feat, e_gibbs, 20, 0.01, 8.)
    # a sampled FP64 audit of one jump in audit_every: its share of the FP64 referee's cost
    audit_every = 8
    rng = np.random.default_rng(held_seed)
    audited = rng.choice(len(phys), size=max(1, len(phys) // audit_every), replace=False)
    fp64_us_per_jump = t_referee * 1e6 / held_out
    audit_us = fp64_us_per_jump * infer_batch / audit_every
    us = s.infer_time_us(infer_batch, reps=20)
    rate = infer_batch * 256 * 256 * 20 / ((us + base_us) * 1e-6)
    rate_refereed = infer_batch * 256 * 256 * 20 / ((us + base_us + phys_us + audit_us) * 1e-6)
    flops, nbytes = forward_flops(infer_batch), forward_bytes(infer_batch, zero_mass=s.zero_mass)
    gbs = nbytes / (us * 1e-6) / 1e9
    roof = dict(bound="hbm", kernel="FNO forward (13 fused kernels: DFT, mixing, inverse + pointwise per layer; BF16 MFMA)",
                achieved=round(gbs, 1), peak=8000.0, unit="GB/s", frac=round(gbs / 8000.0, 4), traffic=None,
                bytes_per_forward=nbytes, flops_per_forward=flops,
                mfma_bf16_tflops=round(flops / (us * 1e-6) / 1e12, 2),
                mfma_bf16_frac=round(flops / (us * 1e-6) / 1e12 / 2500.0, 4),
                note="all products on BF16 MFMA, whose dense peak (2.5 PFLOP/s) is far above what this "
                     "work needs; the bound is the activations' HBM traffic")
    out = dict(pairs=s.npairs, train_instances=len(kept), adam_steps=steps, batch=batch, infer_batch=infer_batch,
               train_s=round(t_train, 1),
               data_s=round(t_data, 1), held_out=len(phys), held_out_unphysical=held_out - len(phys),
               rel_l2_median=float(np.median(err)), rel_l2_mean=float(np.mean(err)), rel_l2_max=float(np.max(err)),
               identity_rel_l2_mean=float(np.mean(ident)), base_steps=base_steps, out_scale=s.out_scale,
               zero_mass=s.zero_mass,
               base_rel_l2_median=float(np.median(base_err)), base_rel_l2_mean=float(np.mean(base_err)),
               base_accepted=int(np.sum(base_err <= 1e-3)), norm_drift_max=float(np.max(drift)),
               accepted=int(ok.sum()), fallback=int((~ok).sum()),
               infer_us_per_batch=round(us, 1), base_us_per_batch=round(base_us, 1),
               value=rate_refereed, value_without_referee=rate, roofline=roof,
               unit="grid-point*steps/s", referee_fp64_s_for_held_out=round(t_referee, 3),
               physical_referee=dict(
                   criteria="norm drift <= 1e-4, energy drift against the relaxation law <= 1e-3, purity <= 1",
                   accepted=int(pok.sum()), false_accepts_l2_above_1e3=int(np.sum(pok & (err > 1e-3))),
                   accepted_and_l2_ok=int(np.sum(pok & (err <= 1e-3))),
                   energy_drift_median=float(np.median(pedrift)), energy_drift_max=float(np.max(pedrift)),
                   fp64_energy_drift_median=float(np.median(ref_edrift)), fp64_energy_drift_max=float(np.max(ref_edrift)),
                   purity_max=float(np.max(ppur)), us_per_batch=round(phys_us, 1),
                   audit_every=audit_every, audited=int(len(audited)),
                   audit_failures=int(np.sum(pok[audited] & (err[audited] > 1e-3))),
                   audit_us_per_batch=round(audit_us, 1)))
    say("[fno] held-out: rel L2 median %.3e (base method %.3e, identity %.3e), accepted %d of %d by the FP64 "
        "referee; %d by the physical referee (%d of them above 1e-3); %.3e pt*steps/s with the referee "
        "(%.3e without)" % (out["rel_l2_median"], out["base_rel_l2_median"], out["identity_rel_l2_mean"],
                            out["accepted"], len(phys), int(pok.sum()), int(np.sum(pok & (err > 1e-3))),
                            rate_refereed, rate))
    return out


if __name__ == "__main__":
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="train and referee the FNO surrogate on propagator trajectories")
    ap.add_argument("--pairs", type=int, default=20000)
    ap.add_argument("--held-out", type=int, default=512)
    ap.add_argument("--epochs", type=float, default=2.0)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--base-steps", type=int, default=2)
    ap.add_argument("--out-scale", type=float, default=1e-2)
    ap.add_argument("--infer-batch", type=int, default=32, help="samples per forward for the timed inference")
    ap.add_argument("--time-inference", type=int, default=0, metavar="REPS",
                    help="only time REPS forward passes of the 256^2 network over --infer-batch samples (random weights)")
    a = ap.parse_args()
    if a.time_inference:
        s = FNOSurrogate(grid_dim=256, width=32, modes=16, layers=4, max_batch=a.infer_batch, seed=0,
                         out_scale=a.out_scale, zero_mass=True)
        us = s.infer_time_us(a.infer_batch, reps=a.time_inference)
        print(json.dumps({"us_per_batch": us, "batch": a.infer_batch, "us_per_sample": us / a.infer_batch, "grid": 256}))
        sys.exit(0)
    print(json.dumps(surrogate_experiment(a.pairs, held_out=a.held_out, epochs=a.epochs, lr=a.lr, infer_batch=a.infer_batch,
                                          base_steps=a.base_steps, out_scale=a.out_scale,
                                          log=lambda m: print(m, file=sys.stderr, flush=True))))
