"""Host logic of the learned adaptive energy grid (negftrans_amd/adaptive.py,
negftrans_amd/workloads.py), CPU only: the RGF call is replaced by a known curve."""
import numpy as np
import pytest

from negftrans_amd import adaptive as A
from negftrans_amd import workloads as W

WIN = 0.02
ED = np.arange(8192) * WIN / 8192
BINS = np.arange(512) * WIN / 512


def curve(E0, w):
    """Two smoothed steps and a narrow resonance (width 3e-5 of a 0.02 window)."""
    return lambda E: (1 / (1 + np.exp(-(E - E0) / w)) + 0.5 / (1 + np.exp(-(E - E0 - 0.004) / (2 * w)))
                      + 0.2 / (1 + ((E - 0.011) / 3e-5) ** 2))


@pytest.fixture
def fake_rgf(monkeypatch):
    state = {}
    monkeypatch.setattr(A, "transmission_at", lambda dev, wl, E: state["f"](np.asarray(E)))
    return state


def test_integrated_is_exact_for_piecewise_linear():
    E = np.array([0.0, 0.3, 0.5, 1.0])
    assert A.integrated(E, 2 * E + 1) == pytest.approx(2.0)


def test_refine_meets_tolerance_within_an_eighth(fake_rgf):
    rng = np.random.default_rng(0)
    for _ in range(6):
        f = curve(rng.uniform(0.003, 0.017), 10 ** rng.uniform(-4.5, -2.5))
        fake_rgf["f"] = f
        G = A.integrated(ED, f(ED))
        E0 = np.linspace(ED[0], ED[-1], 48)
        E, T = A.refine(None, None, E0, f(E0), 1e-4, 1024)
        assert len(E) <= 1024 and np.all(np.diff(E) > 0)
        assert abs(A.integrated(E, T) - G) <= 1e-4


def test_prediction_seeded_mesh_needs_fewer_points(fake_rgf):
    """With an exact prediction the first mesh follows the curvature, and the refinement
    reaches the tolerance with fewer RGF points than from a uniform first mesh of the
    same size; with a prediction that is badly wrong the device falls back."""
    rng = np.random.default_rng(1)
    gains = []
    for _ in range(6):
        f = curve(rng.uniform(0.003, 0.017), 10 ** rng.uniform(-4.5, -2.5))
        fake_rgf["f"] = f
        G = A.integrated(ED, f(ED))
        mu, lv = f(BINS), np.full(512, -9.0)
        Ga, n, fb = A.adaptive_conductance(None, None, mu, lv, BINS, ED[0], ED[-1], 8192, 1e-4, 48)
        assert not fb and abs(Ga - G) <= 1e-4 and n <= 1024
        E0 = np.linspace(ED[0], ED[-1], 48)
        Eb, _ = A.refine(None, None, E0, f(E0), 1e-4, 1024)
        gains.append(len(Eb) / n)
    assert np.median(gains) > 1.5
    Gf, nf, fb = A.adaptive_conductance(None, None, f(BINS) + 1.0, lv, BINS, ED[0], ED[-1], 8192, 1e-4, 48)
    assert fb and nf == 1024


def test_fourier_features_band_limited_closed_form():
    """A potential made of one low-order cosine in x and one in z is reproduced exactly on
    the 16 x 16 x 8 sample grid; a component above the band is dropped."""
    gx, gz = 32, 40
    x = np.arange(gx)[:, None, None] / gx
    z = np.arange(gz)[None, None, :] / gz
    low = np.cos(2 * np.pi * 3 * x) * np.cos(2 * np.pi * 2 * z) * np.ones((gx, gx, gz))
    high = np.cos(2 * np.pi * 11 * x) * np.ones((gx, gx, gz))
    f = W.fourier_features(low + high, scale=1.0).reshape(16, 16, 8)
    xs = np.arange(16)[:, None, None] / 16
    zs = np.arange(8)[None, None, :] / 8
    want = np.cos(2 * np.pi * 3 * xs) * np.cos(2 * np.pi * 2 * zs) * np.ones((16, 16, 8))
    np.testing.assert_allclose(f, want, atol=1e-6)
