"""Writes the golden vectors under tests/golden/ from the CPU oracle.

The reference's own tests hold no golden vectors, known-answer tests or
fixtures for the split-operator path, and its FFTW path cannot run here
(Python 2 sources, pyFFTW/libfftw3 absent). The vectors below therefore come
from the oracle (oracle/fw_oracle.py), which tests/test_oracle.py pins against
the direct DFT, the harmonic closed forms, invariants and planted mistakes
before these files are trusted. They freeze the oracle's output so that a later
change to either the oracle or the HIP path shows as a fixture mismatch.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import cexpr  # noqa: E402
from oracle import fw_oracle as fo  # noqa: E402

V1, K1 = "0.5*X*X+0.1*X*X*X*X", "0.5*P*P"
DV1, DK1 = "X+0.4*X*X*X", "P"


def moyal_quartic(n=256, steps=1000):
    """The quartic Moyal workload at its full size: 256^2, x,p in [-8,8), dt=0.01, 1000 steps;
    series of <x>, <p>, <H>, purity per step and the final W on a 16x16
    subsample (the full W would be 512 KiB)."""
    o = fo.MoyalOracle(n, 8., n, 8., 0.01, V=cexpr.make(V1), K=cexpr.make(K1, args=("P", "t")),
                       diff_V=cexpr.make(DV1), diff_K=cexpr.make(DK1, args=("P", "t")))
    W0 = np.exp(-(o.X - 1.) ** 2 - (o.P - 0.5) ** 2) / np.pi
    o.set_wignerfunction(W0)
    o.propagate(steps)
    frame = o.W[::n // 16, ::n // 16].copy()
    ser = {k: np.array([d[k] for d in o.series]) for k in ("X_average", "P_average", "hamiltonian_average",
                                                          "purity")}
    return dict(ser, W_sub16=frame, full_W_sum=np.sum(o.W), full_W_sq=np.sum(o.W ** 2))


def moyal_small(n=64, steps=200):
    """64x32 rectangular grid with all three masks and a time-dependent V:
    the full final W."""
    cx = dict(X_amplitude=6., P_amplitude=5., dt=0.02, F=0.3)
    V = "0.5*X*X + X*F*sin(t)"
    mxt = "exp(-0.01*dt*X*X*Theta*Theta)"
    mlp = "exp(-0.02*dt*Lambda*Lambda)"
    mxp = "pow(fabs(cos(0.5*M_PI*X/X_amplitude)), dt*0.1)"
    o = fo.MoyalOracle(n, 6., 32, 5., 0.02, V=cexpr.make(V, cx), K=cexpr.make(K1, cx, ("P", "t")),
                       mask_x_theta=cexpr.make(mxt, cx, ("X", "Theta", "t")),
                       mask_lambda_p=cexpr.make(mlp, cx, ("Lambda", "P", "t")),
                       mask_x_p=cexpr.make(mxp, cx, ("X", "P", "t")))
    W0 = np.exp(-(o.X + 0.5) ** 2 / 1.5 - (o.P - 1.0) ** 2 / 0.8)
    o.set_wignerfunction(W0)
    o.propagate(steps)
    return dict(W=o.W.copy(), V=np.array(V), mask_x_theta=np.array(mxt), mask_lambda_p=np.array(mlp),
                mask_x_p=np.array(mxp), F=0.3)


def bloch_double_well(n=64):
    """CUDA-convention Bloch Gibbs state of the double well (V = -x^2/2 + x^4/20) at 64^2."""
    V = "-0.5*X*X + X*X*X*X/20."
    o = fo.BlochOracle(n, 10., n, 10., 0.005, V=cexpr.make(V), K=cexpr.make(K1, args=("P", "t")),
                       kT=0.2, dbeta=0.005, convention="cuda")
    o.get_gibbs_state()
    return dict(W=o.W.copy(), steps=o.steps_done)


def random_collision(n=64, steps=100):
    V = "0.5*X*X + 0.05*X*X*X*X"
    o = fo.RandomCollisionOracle(n, 8., n, 8., 0.01, V=cexpr.make(V), K=cexpr.make(K1, args=("P", "t")),
                                 gamma=0.5, kT=0.5, dbeta=0.01, phase_time_shift=0.0)
    W0 = np.exp(-(o.X - 1.) ** 2 - (o.P - 0.5) ** 2) / np.pi
    o.set_wignerfunction(W0)
    o.propagate(steps)
    return dict(W=o.W.copy(), X_average=np.array([d["X_average"] for d in o.series]))


def rho_vneumann(n=64, steps=100):
    V = "0.5*X*X + 0.05*X*X*X*X"
    o = fo.RhoOracle(n, 8., 0.01, V=cexpr.make(V), K=cexpr.make(K1, args=("P", "t")))
    x = o.Xv
    psi = np.exp(-(x - 1.0) ** 2 / 2 + 0.5j * x)
    o.set_rho(np.outer(psi.conj(), psi))
    o.propagate(steps)
    return dict(rho=o.rho.copy(), wigner=o.get_wignerfunction(), purity=o.get_purity())


def main():
    out = {
        "moyal_quartic_256.npz": moyal_quartic,
        "moyal_masks_64x32.npz": moyal_small,
        "bloch_double_well_64.npz": bloch_double_well,
        "random_collision_64.npz": random_collision,
        "rho_vneumann_64.npz": rho_vneumann,
    }
    for name, fn in out.items():
        d = fn()
        np.savez_compressed(os.path.join(HERE, name), **d)
        print("wrote", name, {k: np.shape(v) for k, v in d.items()})


if __name__ == "__main__":
    main()
