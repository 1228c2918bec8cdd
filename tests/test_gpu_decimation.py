"""Lead terms by decimation (NEGF_LEAD_DECIMATION, extension) through the C-ABI.

k_lead_decimation replaces the reference's closed-form exp(i q Delta) by the surface term
of a semi-infinite flat lead on the device's own lattice.  Bars, each written in its
assertion: relative 1e-10 against the oracle (same decimation restated in numpy) for
transmission, density and LDOS; 1e-9 for the mixed path at refinement tolerance 1e-13;
a V = 0 wire transmits the
number of open lattice channels to 1e-5 (eta = 1e-7).
"""
import dataclasses

import numpy as np
import pytest

from oracle import negf_oracle as O
from tests.test_gpu_parity import RTOL, rel_err, small_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ng():
    from negftrans_amd import negf
    if negf.device_count() < 1:
        pytest.fail("no GPU visible to libnegf_gpu.so")
    return negf


def dec_device(ng, s, prec=0):
    a = s.atomic
    at = ng.Atomic(V_L=a.V_L, V_R=a.V_R, ETA=a.ETA, LX=a.LX, LY=a.LY, LZ=a.LZ, ENCUT=a.ENCUT,
                   MU=a.MU, TEMPERATURE=a.TEMPERATURE, GAP=a.GAP)
    return ng.NegfDevice(s.nx, s.ny, at, s.Vrec, s.NGX, s.NGY, flags=prec | ng.LEAD_DECIMATION)


def test_decimation_transmission_small(ng):
    s, _ = small_system(N_z=9, NK=(2, 1), VL=0.02, VR=-0.01, full=True)
    s = dataclasses.replace(s, lead="decimation")
    E = np.linspace(0.05, 1.2, 17)
    ref = O.transmission(s, E)
    dev = dec_device(ng, s)
    got = dev.transmission(E, (s.kx, s.ky, s.kw))
    assert dev.last_status == 0
    np.testing.assert_allclose(got, ref, rtol=RTOL, atol=1e-12 * max(1.0, np.abs(ref).max()))
    # and it is not the closed-form result
    closed = O.transmission(dataclasses.replace(s, lead="closed-form"), E)
    assert np.max(np.abs(closed - ref)) > 1e-6


def test_decimation_uniform_wire_one_per_channel(ng):
    s, _ = small_system(N_z=7, eta=1e-7)
    s.Vrec[:] = 0
    s = dataclasses.replace(s, lead="decimation")
    a = s.atomic
    delta = a.LZ / s.N_z
    E = np.array([0.03, 0.2, 0.45, 0.8, 1.1])
    got = dec_device(ng, s).transmission(E)
    n_open = [int(np.sum(O.lattice_open(e, 0.0, s.nx, s.ny, a.LX, a.LY, delta, 0.0, 0.0))) for e in E]
    assert max(n_open) >= 5
    np.testing.assert_allclose(got, n_open, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(got, O.transmission(s, E), rtol=RTOL)


def test_decimation_density_and_ldos(ng):
    s, V = small_system(N_z=5, seed=21, NK=(2, 1), eta=1e-4)
    s = dataclasses.replace(s, lead="decimation")
    a = s.atomic
    a.MU, a.TEMPERATURE = 0.6, 300.0
    a.GAP = 6 * O.K_B * a.TEMPERATURE
    minV = float(V.min())
    dev = dec_device(ng, s)
    at = ng.Atomic(V_L=a.V_L, V_R=a.V_R, ETA=a.ETA, LX=a.LX, LY=a.LY, LZ=a.LZ, ENCUT=a.ENCUT,
                   MU=a.MU, TEMPERATURE=a.TEMPERATURE, GAP=a.GAP)
    z, w = ng.density_contour(at, 7, 5, minV)
    got = dev.density_matrix(z, w, (s.kx, s.ky, s.kw))
    assert rel_err(got, np.moveaxis(O.density_matrix(s, 7, 5, minV), 0, 2)) < RTOL
    E = np.array([0.2, 0.7])
    L = dev.ldos_matrix(E, (s.kx, s.ky, s.kw))
    for i, e in enumerate(E):
        assert rel_err(L[:, :, :, i], np.moveaxis(O.ldos_matrix(s, e), 0, 2)) < RTOL


def test_decimation_mixed_path(ng, monkeypatch):
    s, _ = small_system(N_z=40, seed=5)
    s = dataclasses.replace(s, lead="decimation")
    E = np.linspace(0.05, 0.6, 13)
    ref = O.transmission(s, E)
    monkeypatch.setenv("NEGF_MIXED_TOL", "1e-13")
    dev = dec_device(ng, s, prec=ng.PREC_MIXED_FP32)
    monkeypatch.delenv("NEGF_MIXED_TOL")
    got = dev.transmission(E)
    assert dev.last_status == 0
    assert np.all(np.abs(got - ref) <= 1e-9 * (1.0 + np.abs(ref)))


def test_decimation_needs_eta(ng):
    """An open channel at eta = 0 never converges: the call fails with a message (the
    decimation kernel stops after 200 doubling steps)."""
    s, _ = small_system(N_z=4, eta=0.0)
    with pytest.raises(RuntimeError, match="did not converge"):
        dec_device(ng, s).transmission(np.array([0.5]))
