"""Result files (settings / frames / ehrenfest groups of the reference's HDF5
layout, strong_field_physics_cuda_1d.py:110-122,227-237,348-354) and the HHG
spectrum (file_viwer.py:237-260). CPU only."""
import numpy as np
import pytest

from fastwigner_amd.results import ResultsWriter, blackman_window, hhg_spectrum, load_results


class _Dev:
    """Stands in for a device array: only .get()."""

    def __init__(self, a):
        self.a = a

    def get(self):
        return self.a.copy()


class _Prop:
    def __init__(self, rng):
        self.X_wigner = np.linspace(-1, 1, 8)
        self.P_wigner = np.linspace(-2, 2, 8)
        self.t = 0.0
        self.wignerfunction = _Dev(rng.standard_normal((8, 8)))
        self.X_average = list(rng.standard_normal(5))
        self.P_average = list(rng.standard_normal(5))
        self.hamiltonian_average = list(rng.standard_normal(5))
        self.purity_series = []


def test_results_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    p = _Prop(rng)
    path = str(tmp_path / "run.npz")
    settings = dict(X_gridDIM=8, X_amplitude=1.0, dt=0.05, V="-1. / sqrt(X * X + 1.)", E=lambda t: t,
                    omega=0.06, flag=True)
    frames = []
    with ResultsWriter(path, settings, propagator=p) as w:
        for k in range(3):
            p.t = 0.05 * (k + 1) * 10
            p.wignerfunction = _Dev(rng.standard_normal((8, 8)))
            frames.append((p.wignerfunction.a.copy(), p.t))
            assert w.add_propagator_frame(p) == k
        prob = rng.standard_normal(8)
        w.add_frame(np.ones((8, 8)), 2.0, prob=prob)
        w.set_ehrenfest(p, dipole=np.arange(4.0))
    r = load_results(path)
    s = r["settings"]
    assert s["X_gridDIM"] == 8 and s["dt"] == 0.05 and s["V"] == "-1. / sqrt(X * X + 1.)" and s["flag"]
    assert "E" not in s                        # a function is skipped, as h5py would refuse it
    assert np.array_equal(s["X_wigner"], p.X_wigner)
    assert len(r["frames"]) == 4
    for (W, t), f in zip(frames, r["frames"]):
        assert np.array_equal(f["wigner"], W) and f["t"] == t and "prob" not in f
    assert np.array_equal(r["frames"][3]["prob"], prob)
    e = r["ehrenfest"]
    assert np.array_equal(e["X_average"], p.X_average)
    assert np.array_equal(e["hamiltonian_average"], p.hamiltonian_average)
    assert "purity_series" not in e            # empty series are not written
    assert np.array_equal(e["dipole"], np.arange(4.0))
    # the archive needs no pickles
    with np.load(path, allow_pickle=False) as z:
        assert "frames/0/wigner" in z.files


def test_results_failed_run_leaves_no_file(tmp_path):
    path = str(tmp_path / "bad.npz")
    with pytest.raises(RuntimeError):
        with ResultsWriter(path, {"dt": 0.1}) as w:
            w.add_frame(np.zeros((2, 2)), 0.1)
            raise RuntimeError("propagation failed")
    assert not (tmp_path / "bad.npz").exists() and not (tmp_path / "bad.npz.part").exists()


def test_hhg_window_is_numpy_blackman():
    for N in (2, 3, 64, 101, 8378):
        assert np.allclose(blackman_window(N), np.blackman(N), rtol=0, atol=1e-15)
    # two adjacent impulses: the spectrum runs from (w_n + w_n+1)^2 at zero frequency to
    # (w_n - w_n+1)^2 at Nyquist, which pins the window values the spectrum used
    N, n = 64, 20
    J = np.zeros(N)
    J[n] = J[n + 1] = 1.0
    om, s = hhg_spectrum(J, 0.1, 1.0)
    w = np.blackman(N)
    assert abs(s.min() - ((w[n] - w[n + 1]) / (w[n] + w[n + 1])) ** 2) < 1e-14
    assert om[np.argmax(s)] == 0.0


def test_hhg_harmonics_on_bins():
    """Odd harmonics k = 1, 3, 5 of omega placed on FFT bins: peaks at exactly
    k (units of omega), heights in the ratio of the squared amplitudes (the
    window's main-lobe gain is the same for every harmonic; Blackman side
    lobes are below -58 dB), and a symmetric spectrum for a real series."""
    N, dt = 4096, 0.05
    m = 16                                   # omega on bin m
    omega = 2 * np.pi * m / (N * dt)
    t = dt * np.arange(N)
    amps = {1: 1.0, 3: 0.1, 5: 0.01}
    J = sum(a * np.cos(k * omega * t) for k, a in amps.items())
    om, s = hhg_spectrum(J, dt, omega)
    assert np.allclose(s, s[::-1][np.r_[-1, :N - 1]], atol=1e-12)     # s(-w) = s(w) for real J
    for k, a in amps.items():
        i = np.argmin(np.abs(om - k))
        assert abs(om[i] - k) < 1e-12
        assert abs(s[i] / a ** 2 - 1.0) < 1e-4, (k, s[i])
        # the peak is a local maximum
        assert s[i] > s[i - 3] and s[i] > s[i + 3]
    assert abs(s.max() - 1.0) < 1e-15


def test_hhg_rejects_short_series():
    with pytest.raises(ValueError):
        hhg_spectrum([1.0], 0.1, 1.0)
