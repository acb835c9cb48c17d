"""The golden vectors under tests/golden/ against the oracle (CPU).

Regenerates each fixture with the oracle and requires agreement to 1e-13
relative L2 (pocketfft's rounding can differ across builds of scipy in the
last bits). The GPU side of the same fixtures is tests/test_gpu_golden.py.
"""
import os

import numpy as np
import pytest

from tests.golden import make_golden as mg

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(HERE, name), allow_pickle=False)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name,fn", [
    ("moyal_masks_64x32.npz", mg.moyal_small),
    ("bloch_double_well_64.npz", mg.bloch_double_well),
    ("random_collision_64.npz", mg.random_collision),
    ("rho_vneumann_64.npz", mg.rho_vneumann),
    ("moyal_quartic_256.npz", mg.moyal_quartic),
])
def test_oracle_reproduces_golden(name, fn):
    g = load(name)
    d = fn()
    for k in g.files:
        if g[k].dtype.kind in "fc":
            assert rel(d[k], g[k]) <= 1e-13, (name, k)
        else:
            assert np.array_equal(np.asarray(d[k]), g[k]), (name, k)


def test_quartic_256_golden_physics():
    """Quartic Moyal fixture (256^2): normalised, purity conserved by the unitary Moyal
    step (up to the grid's band limit), energy drift O(dt^2)."""
    g = load("moyal_quartic_256.npz")
    pur, H = g["purity"], g["hamiltonian_average"]
    assert abs(g["full_W_sum"] * (16. / 256) ** 2 - 1.0) <= 1e-13
    assert np.max(np.abs(pur - pur[0])) <= 1e-8
    assert np.max(np.abs(H - H[0])) / H[0] <= 1e-3
