"""The HIP propagators against the golden vectors under tests/golden/
(written by tests/golden/make_golden.py from the oracle). FP64 bar: 1e-12
relative L2 on states; observables to 1e-11 relative (<H> weighs the grid
edges by V ~ 4e2, where W's relative rounding differences are largest)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(HERE, name), allow_pickle=False)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def test_quartic_256_series(gpu):
    from fastwigner_amd.wigner_moyal import WignerMoyalHIP1D
    g = load("moyal_quartic_256.npz")
    w = WignerMoyalHIP1D(X_gridDIM=256, X_amplitude=8., P_gridDIM=256, P_amplitude=8., dt=0.01, t=0.,
                         V="0.5*X*X+0.1*X*X*X*X", K="0.5*P*P", diff_V="X+0.4*X*X*X", diff_K="P")
    w.set_wignerfunction(np.exp(-(w.X - 1.) ** 2 - (w.P - 0.5) ** 2) / np.pi)
    W = w.propagate(1000).get()
    assert rel(W[::16, ::16], g["W_sub16"]) <= 1e-12
    for k in ("X_average", "P_average", "hamiltonian_average", "purity"):
        ours = np.array(getattr(w, "purity_series" if k == "purity" else k))
        assert np.max(np.abs(ours - g[k])) <= 1e-11 * max(1.0, np.max(np.abs(g[k]))), k


def test_masks_time_dependent_rectangular(gpu):
    from fastwigner_amd.wigner_moyal import WignerMoyalHIP1D
    g = load("moyal_masks_64x32.npz")
    w = WignerMoyalHIP1D(X_gridDIM=64, X_amplitude=6., P_gridDIM=32, P_amplitude=5., dt=0.02, t=0., F=0.3,
                         V=str(g["V"]), K="0.5*P*P", abs_boundary_x_theta=str(g["mask_x_theta"]),
                         abs_boundary_lambda_p=str(g["mask_lambda_p"]), abs_boundary_x_p=str(g["mask_x_p"]))
    w.set_wignerfunction(np.exp(-(w.X + 0.5) ** 2 / 1.5 - (w.P - 1.0) ** 2 / 0.8))
    assert rel(w.propagate(200).get(), g["W"]) <= 1e-12


def test_bloch_double_well(gpu):
    from fastwigner_amd.wigner_bloch import WignerBlochHIP1D
    g = load("bloch_double_well_64.npz")
    b = WignerBlochHIP1D(X_gridDIM=64, X_amplitude=10., P_gridDIM=64, P_amplitude=10., kT=0.2, dbeta=0.005,
                         t=0., t_initial=0., V="-0.5*X*X + X*X*X*X/20.", K="0.5*P*P")
    assert rel(b.get_gibbs_state().get(), g["W"]) <= 1e-12
    assert b.gibbs_steps == int(g["steps"])


def test_random_collision(gpu):
    from fastwigner_amd.random_collision import WignerRandomCollisionHIP1D
    g = load("random_collision_64.npz")
    w = WignerRandomCollisionHIP1D(X_gridDIM=64, X_amplitude=8., P_gridDIM=64, P_amplitude=8., dt=0.01, t=0.,
                                   gamma=0.5, kT=0.5, dbeta=0.01, V="0.5*X*X + 0.05*X*X*X*X", K="0.5*P*P",
                                   diff_V="X + 0.2*X*X*X", diff_K="P")
    w.set_wignerfunction(np.exp(-(w.X - 1.) ** 2 - (w.P - 0.5) ** 2) / np.pi)
    assert rel(w.propagate(100).get(), g["W"]) <= 1e-12
    assert np.max(np.abs(np.array(w.X_average) - g["X_average"])) <= 1e-12


def test_rho_vneumann(gpu):
    from fastwigner_amd.rho_vneumann import RhoVNeumannHIP1D
    g = load("rho_vneumann_64.npz")
    r = RhoVNeumannHIP1D(X_gridDIM=64, X_amplitude=8., dt=0.01, t=0., V="0.5*X*X + 0.05*X*X*X*X", K="0.5*P*P")
    x = r.X
    psi = np.exp(-(x - 1.0) ** 2 / 2 + 0.5j * x)
    r.set_rho(np.outer(psi.conj(), psi))
    r.propagate(100)
    assert rel(r.rho.get(), g["rho"]) <= 1e-12
    assert rel(r.get_wignerfunction().get(), g["wigner"]) <= 1e-12
    assert abs(r.get_purity() - float(g["purity"])) <= 1e-12
