"""The FP64 transmission autotuner (NEGF_AUTOTUNE, extension) through the C-ABI.

On a small system (N = 21, two k points, N_z = 9): the first call of a job-count class
times the four variants (one- and two-sided, 32-column two-tile and 16-column four-tile
Gauss-Jordan) and returns the fastest's results; every variant's results match the
oracle at relative 1e-10; the next call of the same class reuses the choice and returns
bit-identical results; a call of another class tunes again; each variant forced through
NEGF_TRANS_VARIANT matches the oracle; NEGF_DIAG_ONE_SIDED=1 keeps the one-sided form.
Bars are written in the assertions.
"""
import numpy as np
import pytest

from oracle import negf_oracle as O
from tests.test_gpu_parity import RTOL, small_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ng():
    from negftrans_amd import negf
    if negf.device_count() < 1:
        pytest.fail("no GPU visible to libnegf_gpu.so")
    return negf


def tuned_device(ng, s, flags):
    a = s.atomic
    at = ng.Atomic(V_L=a.V_L, V_R=a.V_R, ETA=a.ETA, LX=a.LX, LY=a.LY, LZ=a.LZ, ENCUT=a.ENCUT)
    return ng.NegfDevice(s.nx, s.ny, at, s.Vrec, s.NGX, s.NGY, flags=flags)


def test_autotune_picks_the_fastest_and_caches_it(ng):
    s, _ = small_system(N_z=9, NK=(2, 1), seed=17)
    E = np.linspace(0.05, 1.2, 12)          # 24 jobs: class 5
    ref = O.transmission(s, E)
    dev = tuned_device(ng, s, ng.AUTOTUNE)
    kp = (s.kx, s.ky, s.kw)
    got = dev.transmission(E, kp)
    secs = dev.tune_seconds()
    assert all(x is not None and x > 0 for x in secs)          # Np = 32: all four variants run
    v = dev.last_variant
    assert v == int(np.argmin(secs))
    np.testing.assert_allclose(got, ref, rtol=RTOL, atol=1e-12)
    again = dev.transmission(E, kp)
    assert dev.last_variant == v and dev.tune_seconds() == secs  # cached: no new tuning
    assert np.array_equal(again, got)
    dev.transmission(E[:3], kp)             # 6 jobs: another class, tuned afresh
    assert dev.tune_seconds() != secs
    rule = tuned_device(ng, s, 0)            # the fixed rule: two workgroups per job fit
    np.testing.assert_allclose(rule.transmission(E, kp), ref, rtol=RTOL, atol=1e-12)
    assert rule.last_variant == 1


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_every_variant_matches_the_oracle(ng, variant, monkeypatch):
    """Each variant forced (NEGF_TRANS_VARIANT) against the oracle at relative 1e-10: the
    variants differ only in how the same sweep is tiled and cut."""
    s, _ = small_system(N_z=9, NK=(2, 1), seed=17)
    E = np.linspace(0.05, 1.2, 12)
    ref = O.transmission(s, E)
    monkeypatch.setenv("NEGF_TRANS_VARIANT", str(variant))
    dev = tuned_device(ng, s, 0)
    monkeypatch.delenv("NEGF_TRANS_VARIANT")
    got = dev.transmission(E, (s.kx, s.ky, s.kw))
    assert dev.last_variant == variant and dev.last_status == 0
    np.testing.assert_allclose(got, ref, rtol=RTOL, atol=1e-12)


def test_one_sided_switch_bypasses_tuning(ng, monkeypatch):
    s, _ = small_system(N_z=9, seed=5)
    monkeypatch.setenv("NEGF_DIAG_ONE_SIDED", "1")
    dev = tuned_device(ng, s, ng.AUTOTUNE)
    monkeypatch.delenv("NEGF_DIAG_ONE_SIDED")
    dev.transmission(np.linspace(0.05, 1.2, 8))
    assert dev.last_variant == 0 and dev.tune_seconds() == [None] * 4
