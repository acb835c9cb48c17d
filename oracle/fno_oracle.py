"""numpy restatement of the Fourier-neural-operator surrogate (forward, loss,
backward) -- TEST INFRASTRUCTURE ONLY: the checker of the HIP kernels in
fastwigner_amd/csrc/fw_fno.cpp, never called by the product.

The reference has no learned component; this network is defined by the build
(the surrogate of k split-operator steps). Activations are channel-first,
v[b][c][p][x]. The truncated transforms are written as products with explicit
cosine/sine matrices, exactly as the HIP kernels compute them:

  forward DFT of a channel image (p, x):
    T[p][r] = sum_x v[p][x] Fx[r][x]          Fx = [cos(2 pi ky x/N) ; -sin(2 pi ky x/N)], ky < M
    O[s][r] = sum_p Fp[s][p] T[p][r]          Fp = [cos(2 pi kp p/N) ; sin(2 pi kp p/N)], 2M values of kp
    Z[i][ky] = (O[i][ky] + O[2M+i][M+ky]) + i (O[i][M+ky] - O[2M+i][ky])   (= sum v e^{-i theta})
  inverse (numpy irfft2 of the truncated spectrum Y):
    H[p][ky] = sum_i Y[i][ky] e^{+2 pi i kp_i p/N}     (Hr, Hi through the same Fp)
    S[p][x]  = sum_ky Gx[ky][x] Hr[p][ky] + Gx[M+ky][x] Hi[p][ky],  Gx = (c_ky/N^2) Fx, c_0 = 1, c_ky = 2

With ``bf16=True`` it rounds to bfloat16 (round to nearest even, through
float32) exactly what the HIP kernels round: the stored activations (v, the
lift output, dpre), the transform matrices, the intermediates T and H of both
transforms, the spectra Z / Y entering the mode mixing and the inverse, the
mixing and pointwise weights, forward and backward. Its gradient treats each
rounding as the identity. GELU uses scipy's erf; the kernels evaluate erf by
Abramowitz & Stegun 7.1.26, within 4.3e-7 of this GELU in FP32 (far below the
bf16 resolution of the stored activations). With ``bf16=False`` (no rounding) the network is the
exact truncated-FFT network and the gradient the exact one: pinned in
tests/test_fno_oracle.py by central finite differences, by numpy's rfft2 /
irfft2, and by a direct DFT sum for identity mode weights.

Parameter vector (float32 on the GPU, float64 here), in this order:
  Pw[8][C], pb[C], for each layer {R[2M*M][C][C][2], Wp[C][C], b[C]}, Q[C], qb[1]
Features: f = [W, x/X_amplitude, p/P_amplitude, a2, a3, a4, gamma, kT].
"""
import numpy as np
from scipy.special import erf


def bf16(x):
    """Round to bfloat16 (nearest even) via float32; returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


class FNOOracle:
    def __init__(self, N=256, C=32, M=16, L=4, bf16=True, out_scale=1.0, zero_mass=False):
        self.N, self.C, self.M, self.L, self.use_bf16 = N, C, M, L, bf16
        # out = W + out_scale c, c = v_L Q + qb; with zero_mass, c - mean(c) per sample
        self.out_scale, self.zero_mass = float(out_scale), bool(zero_mass)
        self.nm = 2 * M * M
        o = 0
        self.o_Pw = o; o += 8 * C
        self.o_pb = o; o += C
        self.o_R, self.o_Wp, self.o_b = [], [], []
        for _ in range(L):
            self.o_R.append(o); o += self.nm * C * C * 2
            self.o_Wp.append(o); o += C * C
            self.o_b.append(o); o += C
        self.o_Q = o; o += C
        self.o_qb = o; o += 1
        self.nparam = o
        self.kp = np.concatenate([np.arange(M), np.arange(N - M, N)])
        # transform matrices; angles reduced exactly in integers, (k n) mod N
        x = np.arange(N)
        a = 2.0 * np.pi * ((np.arange(M)[:, None] * x[None, :]) % N) / N
        self.Fx = self._r(np.concatenate([np.cos(a), -np.sin(a)]))        # (2M, N)
        b = 2.0 * np.pi * ((self.kp[:, None] * x[None, :]) % N) / N
        self.Fp = self._r(np.concatenate([np.cos(b), np.sin(b)]))         # (4M, N)
        # irfft weights c_ky / N^2 (powers of two: scaling a rounded matrix by them is exact)
        self.cky = np.where(np.arange(M) == 0, 1.0, 2.0) / (N * N)
        self.Gx = self.Fx * np.concatenate([self.cky, self.cky])[:, None]

    # -- helpers -----------------------------------------------------------------------
    def _r(self, x):
        return bf16(x) if self.use_bf16 else x

    def unpack(self, P):
        C, nm = self.C, self.nm
        d = dict(Pw=P[self.o_Pw:self.o_Pw + 8 * C].reshape(8, C), pb=P[self.o_pb:self.o_pb + C],
                 Q=P[self.o_Q:self.o_Q + C], qb=P[self.o_qb])
        d["R"] = [P[o:o + nm * C * C * 2].reshape(nm, C, C, 2) for o in self.o_R]
        d["Wp"] = [P[o:o + C * C].reshape(C, C) for o in self.o_Wp]
        d["b"] = [P[o:o + C] for o in self.o_b]
        return d

    def features(self, W, feat):
        """f[b][p][x][j]; x/X_amplitude = 2 (x_index - N/2) / N, likewise p."""
        B, N = W.shape[0], self.N
        g = (np.arange(N) - 0.5 * N) * (2.0 / N)
        f = np.empty((B, N, N, 8))
        f[..., 0] = W
        f[..., 1] = g[np.newaxis, np.newaxis, :]
        f[..., 2] = g[np.newaxis, :, np.newaxis]
        for j in range(5):
            f[..., 3 + j] = np.asarray(feat)[:, j][:, np.newaxis, np.newaxis]
        return f

    def modes(self, v):
        """Z[b, m = i*M + ky, c] (complex) of channel images v[b][c][p][x] on the retained modes."""
        B, C, M = v.shape[0], v.shape[1], self.M
        T = self._r(np.einsum("bcpx,rx->bcpr", v, self.Fx, optimize=True))
        O = np.einsum("sp,bcpr->bcsr", self.Fp, T, optimize=True)
        Zr = O[:, :, :2 * M, :M] + O[:, :, 2 * M:, M:]
        Zi = O[:, :, :2 * M, M:] - O[:, :, 2 * M:, :M]
        return (Zr + 1j * Zi).reshape(B, C, self.nm).transpose(0, 2, 1)

    def inverse(self, Y, G=None):
        """S[b][o][p][x] = sum over the retained modes of Y[b, m, o] (rounded) through the
        p transform (Fp) and the x table G (default: the irfft weights Gx)."""
        G = self.Gx if G is None else G
        B, O, M = Y.shape[0], Y.shape[-1], self.M
        Y = Y.transpose(0, 2, 1).reshape(B, O, 2 * M, M)
        Yr, Yi = self._r(Y.real), self._r(Y.imag)
        cs, sn = self.Fp[:2 * M], self.Fp[2 * M:]
        Hr = np.einsum("ip,boik->bopk", cs, Yr, optimize=True) - np.einsum("ip,boik->bopk", sn, Yi, optimize=True)
        Hi = np.einsum("ip,boik->bopk", sn, Yr, optimize=True) + np.einsum("ip,boik->bopk", cs, Yi, optimize=True)
        Hr, Hi = self._r(Hr), self._r(Hi)
        return (np.einsum("kx,bopk->bopx", G[:M], Hr, optimize=True)
                + np.einsum("kx,bopk->bopx", G[M:], Hi, optimize=True))

    def mix(self, Z, R):
        """Y[b, m, o] = sum_c Z[b, m, c] R[m, c, o] with the real 2C x 2O block
        [[Rr, Ri], [-Ri, Rr]] and bf16-rounded operands."""
        Zr, Zi = self._r(Z.real), self._r(Z.imag)
        Rr, Ri = self._r(R[..., 0]), self._r(R[..., 1])
        Yr = np.einsum("bmc,mco->bmo", Zr, Rr, optimize=True) - np.einsum("bmc,mco->bmo", Zi, Ri, optimize=True)
        Yi = np.einsum("bmc,mco->bmo", Zr, Ri, optimize=True) + np.einsum("bmc,mco->bmo", Zi, Rr, optimize=True)
        return Yr + 1j * Yi

    # -- forward -------------------------------------------------------------------------
    def forward(self, P, W, feat, keep=False):
        p = self.unpack(P)
        f = self.features(W, feat)
        v = [self._r(np.einsum("bpxj,jc->bcpx", f, p["Pw"], optimize=True) + p["pb"][None, :, None, None])]
        pres, Zs = [], []
        for l in range(self.L):
            Z = self.modes(v[l])
            Zs.append(Z)
            S = self.inverse(self.mix(Z, p["R"][l]))
            Pt = np.einsum("bcpx,co->bopx", v[l], self._r(p["Wp"][l]), optimize=True)
            pre = S + Pt + p["b"][l][None, :, None, None]
            pres.append(pre)
            v.append(self._r(gelu(pre) if l < self.L - 1 else pre))
        c = np.einsum("bcpx,c->bpx", v[-1], p["Q"], optimize=True) + p["qb"]
        if self.zero_mass:
            c = c - c.mean(axis=(1, 2), keepdims=True)
        out = W + self.out_scale * c
        if keep:
            return out, dict(f=f, v=v, pre=pres, Z=Zs, p=p)
        return out

    def loss(self, out, target):
        se = np.sum((out - target) ** 2, axis=(1, 2))
        st = np.sum(target ** 2, axis=(1, 2))
        return float(np.mean(se / st)), st

    # -- backward ------------------------------------------------------------------------
    def loss_grad(self, P, W, feat, target):
        out, k = self.forward(P, W, feat, keep=True)
        L, st = self.loss(out, target)
        B, C, M = W.shape[0], self.C, self.M
        p = k["p"]
        dout = 2.0 * (out - target) / st[:, None, None] / B
        g = np.zeros(self.nparam)
        dout = self.out_scale * dout     # from here on, the gradient w.r.t. the projection's output c
        if self.zero_mass:
            dout = dout - dout.mean(axis=(1, 2), keepdims=True)
        g[self.o_Q:self.o_Q + C] = np.einsum("bcpx,bpx->c", k["v"][-1], dout, optimize=True)
        g[self.o_qb] = dout.sum()
        dv = p["Q"][None, :, None, None] * dout[:, None]
        # dY = (c_ky / N^2) Z(dpre): the adjoint of the inverse is the forward transform
        # (Gx^T = (c_ky/N^2) Fx, Gp^T = Fp), scaled per ky
        yscale = np.tile(self.cky, 2 * M)
        for l in range(self.L - 1, -1, -1):
            dpre = dv * gelu_d(k["pre"][l]) if l < self.L - 1 else dv
            g[self.o_b[l]:self.o_b[l] + C] = dpre.sum(axis=(0, 2, 3))
            d16 = self._r(dpre)
            Wp = self._r(p["Wp"][l])
            dvn = np.einsum("bopx,co->bcpx", d16, Wp, optimize=True)
            g[self.o_Wp[l]:self.o_Wp[l] + C * C] = np.einsum("bcpx,bopx->co", k["v"][l], d16, optimize=True).ravel()
            dY = self.modes(d16) * yscale[None, :, None]
            dYr, dYi = self._r(dY.real), self._r(dY.imag)
            R = p["R"][l]
            Rr, Ri = self._r(R[..., 0]), self._r(R[..., 1])
            # dZr = dYr Rr^T + dYi Ri^T ; dZi = -dYr Ri^T + dYi Rr^T  (transpose of [[Rr, Ri], [-Ri, Rr]])
            dZr = (np.einsum("bmo,mco->bmc", dYr, Rr, optimize=True)
                   + np.einsum("bmo,mco->bmc", dYi, Ri, optimize=True))
            dZi = (-np.einsum("bmo,mco->bmc", dYr, Ri, optimize=True)
                   + np.einsum("bmo,mco->bmc", dYi, Rr, optimize=True))
            Zr, Zi = self._r(k["Z"][l].real), self._r(k["Z"][l].imag)
            # dW = [Zr; Zi]^T [dYr, dYi]; dRr = dW11 + dW22, dRi = dW12 - dW21
            dRr = (np.einsum("bmc,bmo->mco", Zr, dYr, optimize=True)
                   + np.einsum("bmc,bmo->mco", Zi, dYi, optimize=True))
            dRi = (np.einsum("bmc,bmo->mco", Zr, dYi, optimize=True)
                   - np.einsum("bmc,bmo->mco", Zi, dYr, optimize=True))
            gR = np.stack([dRr, dRi], axis=-1)
            g[self.o_R[l]:self.o_R[l] + gR.size] = gR.ravel()
            # adjoint of the forward transform: the inverse's structure with the table Fx
            # (unit weights): dv = Re sum_modes (dZr + i dZi) e^{+i theta}
            dv = dvn + self.inverse(dZr + 1j * dZi, G=self.Fx)
        f = k["f"]
        g[self.o_Pw:self.o_Pw + 8 * C] = np.einsum("bpxj,bcpx->jc", f, dv, optimize=True).ravel()
        g[self.o_pb:self.o_pb + C] = dv.sum(axis=(0, 2, 3))
        return L, g


def init_params(o, seed=0, spectral_scale=None):
    """Initial parameters: lift and pointwise ~ N(0, 1/fan_in), spectral ~ U(-s, s)
    with s = 1/(C*C) (as the original FNO), projection zero (the untrained network is
    the identity map W -> W)."""
    rng = np.random.default_rng(seed)
    C = o.C
    P = np.zeros(o.nparam)
    P[o.o_Pw:o.o_Pw + 8 * C] = rng.standard_normal(8 * C) / np.sqrt(8)
    s = spectral_scale if spectral_scale is not None else 1.0 / (C * C)
    for l in range(o.L):
        P[o.o_R[l]:o.o_R[l] + o.nm * C * C * 2] = rng.uniform(-s, s, o.nm * C * C * 2)
        P[o.o_Wp[l]:o.o_Wp[l] + C * C] = rng.standard_normal(C * C) / np.sqrt(C)
    return P
