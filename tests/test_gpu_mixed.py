"""Mixed-precision transmission (NEGF_PREC_MIXED_FP32) through the C-ABI.

The sweep runs in FP32 and keeps every generator in HBM.  G_{N_z,1} on the open
channels is then refined in FP64 against the exact block-tridiagonal operator until the
estimated error of T is below the context's tolerance (NEGF_MIXED_TOL, default 1e-7).
Bars, each written in its assertion:

- with the tolerance at 1e-13, against the CPU oracle on small systems: |dT| <= 1e-9 (1 + |T|)
  (the refinement converges to the FP64 answer);
- with the default tolerance: max |dT| <= 1e-6 against the FP64 GPU path and the oracle,
  the accuracy gate of the mixed-vs-FP64 comparison, on the c0 geometry and at the metric
  size (N = 317, N_z = 1000).
"""
import numpy as np
import pytest

from oracle import negf_oracle as O
from tests.test_gpu_parity import device_for, oracle_system, small_system  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ng():
    from negftrans_amd import negf
    if negf.device_count() < 1:
        pytest.fail("no GPU visible to libnegf_gpu.so")
    return negf


def mixed_device(ng, s, monkeypatch=None, tol=None):
    """A mixed-precision context; tol, if given, is its refinement tolerance on T (read
    from NEGF_MIXED_TOL when the context is created)."""
    if tol is not None:
        monkeypatch.setenv("NEGF_MIXED_TOL", repr(tol))
    a = s.atomic
    at = ng.Atomic(V_L=a.V_L, V_R=a.V_R, ETA=a.ETA, LX=a.LX, LY=a.LY, LZ=a.LZ, ENCUT=a.ENCUT)
    dev = ng.NegfDevice(s.nx, s.ny, at, s.Vrec, s.NGX, s.NGY, flags=ng.PREC_MIXED_FP32)
    if tol is not None:
        monkeypatch.delenv("NEGF_MIXED_TOL")
    return dev


@pytest.mark.parametrize("N_z,seed", [(9, 3), (40, 5), (2, 9)])
def test_mixed_small_vs_oracle(ng, N_z, seed, monkeypatch):
    s, _ = small_system(N_z=N_z, seed=seed)
    kin = O.kinetic(s.nx, s.ny, s.atomic.LX, s.atomic.LY, 0.0, 0.0)
    E = np.linspace(0.05, 0.6, 13)
    assert max(np.sum(2 * e - kin > 0) for e in E) <= 16
    ref = O.transmission(s, E)
    dev = mixed_device(ng, s, monkeypatch, tol=1e-13)
    got = dev.transmission(E)
    assert dev.last_status == 0
    assert np.all(np.abs(got - ref) <= 1e-9 * (1.0 + np.abs(ref)))
    dev = mixed_device(ng, s)
    got = dev.transmission(E)
    assert dev.last_status == 0
    assert np.all(np.abs(got - ref) <= 1e-6)


@pytest.mark.parametrize("R2,seed", [(40.5, 1), (60.5, 2), (70.5, 4), (90.5, 6)])
def test_mixed_wave_counts_vs_oracle(ng, R2, seed, monkeypatch):
    """Block sizes whose waves sit unevenly on the four SIMDs: Np = 160, 192, 224 and 320
    (5, 6, 7 and 10 waves), where the trailing update hands column tiles from one or two
    donor waves to helper waves.  Against the CPU oracle at tolerance 1e-13."""
    s, _ = small_system(N_z=6, seed=seed, R2=R2)
    kin = O.kinetic(s.nx, s.ny, s.atomic.LX, s.atomic.LY, 0.0, 0.0)
    E = np.linspace(0.05, 0.6, 5)
    assert max(np.sum(2 * e - kin > 0) for e in E) <= 16
    ref = O.transmission(s, E)
    dev = mixed_device(ng, s, monkeypatch, tol=1e-13)
    got = dev.transmission(E)
    assert dev.last_status == 0
    assert np.all(np.abs(got - ref) <= 1e-9 * (1.0 + np.abs(ref)))


def test_mixed_qpc_c0_vs_fp64(ng, monkeypatch):
    """Config 0 geometry (N = 61, N_z = 200): mixed against the FP64 GPU path, within the
    1e-6 gate at the default tolerance and within 1e-9 at tolerance 1e-13."""
    from negftrans_amd import workloads
    wl = workloads.qpc_c0()
    NGX, NGY = wl.ngx()
    vrec = ng.local_potential_real_to_planewave(wl.V_real_eV / O.HARTREE, NGX, NGY)
    E = wl.energies_hartree()[::17][:24]
    f64 = ng.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY).transmission(E)
    mix = ng.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY, flags=ng.PREC_MIXED_FP32).transmission(E)
    assert np.max(np.abs(mix - f64)) <= 1e-6
    monkeypatch.setenv("NEGF_MIXED_TOL", "1e-13")
    dev = ng.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY, flags=ng.PREC_MIXED_FP32)
    monkeypatch.delenv("NEGF_MIXED_TOL")
    mix = dev.transmission(E)
    assert np.max(np.abs(mix - f64)) <= 1e-9


def test_mixed_disorder_c1_gate(ng):
    """Metric configuration (N = 317, N_z = 1000), 32 energies spread over the 0-30 meV
    grid: max |T_mixed - T_FP64| <= 1e-6."""
    from negftrans_amd import workloads
    wl = workloads.disorder_c1()
    NGX, NGY = wl.ngx()
    vrec = ng.local_potential_real_to_planewave(wl.V_real_eV / O.HARTREE, NGX, NGY)
    E = wl.energies_hartree()[5::128]
    f64 = ng.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY).transmission(E)
    dev = ng.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY, flags=ng.PREC_MIXED_FP32)
    mix = dev.transmission(E)
    assert dev.last_status == 0
    dT = np.max(np.abs(mix - f64))
    print(f"max |dT| mixed vs FP64 over {len(E)} energies: {dT:.3e}")
    assert dT <= 1e-6


def test_mixed_disorder_c1_256_energies_vs_oracle(ng):
    """The mixed path as the benchmark runs it (N = 317, N_z = 1000, 256 energies in one
    call, workgroups taking jobs in turn, default tolerance): every energy within the 1e-6
    gate of the FP64 GPU path, and two energies within it of the oracle."""
    from negftrans_amd import workloads
    wl = workloads.disorder_c1()
    NGX, NGY = wl.ngx()
    vrec = ng.local_potential_real_to_planewave(wl.V_real_eV / O.HARTREE, NGX, NGY)
    E = wl.energies_hartree()[(np.arange(256) * 16 + 5) % 4096]
    dev = ng.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY, flags=ng.PREC_MIXED_FP32)
    mix = dev.transmission(E)
    assert dev.last_status == 0
    f64 = ng.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY).transmission(E)
    assert np.max(np.abs(mix - f64)) <= 1e-6
    s = oracle_system(wl, np.asarray(vrec), NGX, NGY)
    for i in (40, 201):
        ref = O.transmission_tau(s, E[i], 0, faithful_solver=False)
        assert abs(mix[i] - ref) <= 1e-6


def test_mixed_rejects_more_than_16_open_channels(ng):
    s, _ = small_system(N_z=4, seed=3)
    kin = O.kinetic(s.nx, s.ny, s.atomic.LX, s.atomic.LY, 0.0, 0.0)
    E = np.array([1.5])
    assert np.sum(2 * E[0] - kin > 0) > 16
    with pytest.raises(RuntimeError, match="open channels"):
        mixed_device(ng, s).transmission(E)
