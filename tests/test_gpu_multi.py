"""Uneven batches (negf_multi, extension) through the C-ABI.

Several devices with different N and N_z in one hipGraph.  Bars, each written in its
assertion:
- with the matrix-core variant forced, every device's T equals its own single-device
  negf_transmission_batch result bit for bit (same kernel, same arithmetic);
- with the VALU variant forced, and with the autotuned choice, every device matches the
  CPU oracle to relative 1e-10;
- graph replays after a potential update equal eager single-device runs with the new
  potential, on every replay;
- LDOS at selected slices (negf_multi_ldos) equals each device's negf_ldos_batch at those
  slices to 1e-12 relative (k-point sums by atomics) and the oracle to 1e-10.
"""
import os

import numpy as np
import pytest

from oracle import negf_oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-10


@pytest.fixture(scope="module")
def ng():
    from negftrans_amd import negf
    if negf.device_count() < 1:
        pytest.fail("no GPU visible to libnegf_gpu.so")
    return negf


@pytest.fixture
def variant():
    """Set NEGF_MULTI_VARIANT for one test and restore it afterwards."""
    old = os.environ.get("NEGF_MULTI_VARIANT")

    def set_(v):
        if v is None:
            os.environ.pop("NEGF_MULTI_VARIANT", None)
        else:
            os.environ["NEGF_MULTI_VARIANT"] = v
    yield set_
    set_(old)


def system(R2, N_z, seed, L=12.0, NK=(1, 1), eta=1e-6, amp=0.02):
    ENCUT = R2 * (2 * np.pi / L) ** 2 / 2
    nx, ny = O.planewave_basis(ENCUT, L, L)
    NGX = NGY = 2 * int(np.ceil(np.sqrt(R2))) + 1
    N_x = N_y = 2 * NGX + 4
    V = amp * np.random.default_rng(seed).standard_normal((N_x, N_y, N_z)) + 0.01
    Vrec = O.v_reciprocal_all(V, NGX, NGY)
    a = O.Atomic(ETA=eta, LX=L, LY=L, LZ=N_z * 0.7, ENCUT=ENCUT)
    kx, ky, kw = O.kpoint_mesh(NK[0], NK[1], L, L)
    return O.System(a, nx, ny, Vrec, NGX, NGY, kx, ky, kw)


@pytest.fixture(autouse=True)
def one_sided_references(monkeypatch):
    """Single-device reference runs use the one-workgroup sweep, the form the batch runs
    (a call with few jobs otherwise takes the two-workgroup form, whose summation order
    differs; test_gpu_parity checks that form against the oracle and this one)."""
    monkeypatch.setenv("NEGF_DIAG_ONE_SIDED", "1")


def device(ng, s):
    a = s.atomic
    at = ng.Atomic(V_L=a.V_L, V_R=a.V_R, ETA=a.ETA, LX=a.LX, LY=a.LY, LZ=a.LZ, ENCUT=a.ENCUT)
    return ng.NegfDevice(s.nx, s.ny, at, s.Vrec, s.NGX, s.NGY)


# four devices, three padded sizes (32, 32, 64, 128), N_z from 3 to 14
SMALL = [dict(R2=6.5, N_z=9, seed=1), dict(R2=4.5, N_z=14, seed=2, NK=(2, 1)),
         dict(R2=12.5, N_z=3, seed=3), dict(R2=30.5, N_z=6, seed=4)]


def small_set(ng):
    systems = [system(**p) for p in SMALL]
    devs = [device(ng, s) for s in systems]
    energies = [np.linspace(0.05, 1.1, 7 + 2 * i) for i in range(len(systems))]
    kps = [(s.kx, s.ky, s.kw) for s in systems]
    return systems, devs, energies, kps


def test_multi_mfma_equals_single_device_runs(ng, variant):
    variant("mfma")
    systems, devs, energies, kps = small_set(ng)
    assert len({d.N for d in devs}) == 4
    got = ng.Multi(devs).transmission(energies, kps)
    for d, E, kp, g in zip(devs, energies, kps, got):
        ref = d.transmission(E, kp)
        assert np.array_equal(g, ref), (d.N, np.max(np.abs(g - ref)))


@pytest.mark.parametrize("v", ["valu", None])
def test_multi_vs_oracle(ng, variant, v):
    variant(v)
    systems, devs, energies, kps = small_set(ng)
    m = ng.Multi(devs)
    got = m.transmission(energies, kps)
    assert m.last_status == 0 and m.last_seconds > 0
    for s, E, g in zip(systems, energies, got):
        ref = O.transmission(s, E)
        np.testing.assert_allclose(g, ref, rtol=RTOL, atol=1e-12 * max(1.0, np.abs(ref).max()))
    if v is None:
        for d in devs:
            assert ng.Multi.variant(-(-d.N // 32) * 32) in ("mfma", "valu")


def test_multi_replay_after_potential_update(ng, variant):
    variant("mfma")
    systems, devs, energies, kps = small_set(ng)
    m = ng.Multi(devs)
    first = m.transmission(energies, kps)
    rng = np.random.default_rng(99)
    for rep in range(3):
        # eager work on the GPU between replays: a new potential for device 1 and a
        # single-device run of device 0
        s = systems[1]
        N_x = 2 * s.NGX + 4
        V = 0.02 * rng.standard_normal((N_x, N_x, s.N_z)) + 0.01
        devs[1].set_potential(O.v_reciprocal_all(V, s.NGX, s.NGY))
        eager0 = devs[0].transmission(energies[0], kps[0])
        got = m.replay()
        assert m.last_status == 0
        assert np.array_equal(got[0], first[0]) and np.array_equal(got[0], eager0)
        assert np.array_equal(got[1], devs[1].transmission(energies[1], kps[1]))
        assert not np.array_equal(got[1], first[1])
        for i in (2, 3):
            assert np.array_equal(got[i], first[i])


def test_multi_ldos_selected_slices(ng):
    """LDOS blocks at three slices per device in one graph, against each device's own
    negf_ldos_batch (relative 1e-12: the k-point sums use atomics) and, for one device,
    against the oracle (1e-10)."""
    systems, devs, energies, kps = small_set(ng)
    sl = [[0, s.N_z // 2, s.N_z - 1] for s in systems]
    E = [e[:4] for e in energies]
    m = ng.Multi(devs)
    got = m.ldos(E, sl, kps)
    assert m.last_status == 0
    for d, e, g, kp, slc in zip(devs, E, got, kps, sl):
        ref = d.ldos_matrix(e, kp)[:, :, slc, :]
        assert np.max(np.abs(g - ref)) <= 1e-12 * np.abs(ref).max()
    ref0 = np.moveaxis(O.ldos_matrix(systems[0], E[0][0]), 0, 2)[:, :, sl[0]]
    assert np.linalg.norm(got[0][:, :, :, 0] - ref0) <= RTOL * np.linalg.norm(ref0)


def test_multi_large_block_valu_path(ng):
    """N > 512 (padded 16-wide, one thread per row, the only variant there) next to a
    small device, against the oracle."""
    big = system(R2=170.5, N_z=3, seed=6, eta=1e-5)
    assert big.N > 512
    small = system(R2=6.5, N_z=5, seed=7)
    devs = [device(ng, big), device(ng, small)]
    E = [np.array([0.3, 0.9]), np.array([0.2, 0.6, 1.0])]
    got = ng.Multi(devs).transmission(E)
    for s, e, g in zip((big, small), E, got):
        ref = O.transmission(s, e)
        np.testing.assert_allclose(g, ref, rtol=RTOL, atol=1e-12 * max(1.0, np.abs(ref).max()))
    single = devs[0].transmission(E[0])
    np.testing.assert_allclose(single, O.transmission(big, E[0]), rtol=RTOL)


def test_multi_uneven_block_sizes_vs_single(ng, variant):
    """The uneven-batch block sizes N in {61, 113, 197, 317, 509} with N_z from 40 to
    300, 12 energies each, autotuned: every device matches its single-device run to
    relative 1e-10, and bit for bit where the matrix-core variant was chosen (the
    references run the one-workgroup form, as the batch does)."""
    variant(None)
    from negftrans_amd import workloads
    params = [(61, 300), (113, 40), (197, 160), (317, 80), (509, 60)]
    systems, devs, energies = [], [], []
    for i, (N, N_z) in enumerate(params):
        R2 = {61: 18.5, 113: 36.5, 197: 64.5, 317: 100.5, 509: 162.5}[N]
        s = system(R2=R2, N_z=N_z, seed=10 + i, L=200.0, amp=2e-5, eta=1e-8)
        assert s.N == N
        systems.append(s)
        devs.append(device(ng, s))
        kin_max = O.kinetic(s.nx, s.ny, s.atomic.LX, s.atomic.LY, 0, 0).max()
        energies.append(np.linspace(0.0005, 0.3 * kin_max, 12))
    m = ng.Multi(devs)
    got = m.transmission(energies)
    assert m.last_status == 0
    for d, E, g in zip(devs, energies, got):
        ref = d.transmission(E)
        np.testing.assert_allclose(g, ref, rtol=RTOL, atol=1e-13 * max(1.0, np.abs(ref).max()))
        Np = -(-d.N // 32) * 32
        if ng.Multi.variant(Np) == "mfma":
            assert np.array_equal(g, ref)
    # LDOS at 8 slices per device, two energies, against each device's own LDOS
    sl = [np.linspace(0, d.N_z - 1, 8).round().astype(int) for d in devs]
    E2 = [e[3:5] for e in energies]
    L = m.ldos(E2, sl)
    assert m.last_status == 0
    for d, e, g, slc in zip(devs, E2, L, sl):
        ref = d.ldos_matrix(e)[:, :, slc, :]
        assert np.max(np.abs(g - ref)) <= 1e-12 * np.abs(ref).max()


def test_uneven_c2_workload(ng, variant):
    """The uneven-batch configuration itself: the 48 devices of workloads.uneven_c2 (N in
    {61, 113, 197, 317, 509}, N_z from 228 to 1451, seeded disorder, m* = 0.067), all in
    one graph.  64 of the configuration's 256 energies per device (a strided quarter of
    its 0-30 meV grid), to keep the test near a minute; the benchmark's uneven leg times
    the full batch.  Checks, each in its assertion:
    - 0 <= T <= number of open channels on every device and energy;
    - the N = 61 device (N_z = 228) and the N = 509 device with the shortest wire
      (N_z = 305) against the oracle at relative 1e-10 on two energies each;
    - LDOS blocks at 8 slices per device (two energies): the spectral part
      (L - L^dagger) / 2i of each block is positive semidefinite (L = -G / pi, retarded),
      and two devices equal their own single-device negf_ldos_batch at those slices."""
    variant(None)
    from negftrans_amd import workloads
    wls = workloads.uneven_c2()
    devs, Es, vrecs = [], [], []
    for wl in wls:
        NGX, NGY = wl.ngx()
        vrec = ng.local_potential_real_to_planewave(wl.V_real_eV / O.HARTREE, NGX, NGY)
        vrecs.append(np.asarray(vrec))
        devs.append(ng.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY))
        Es.append(wl.energies_hartree()[1::4])
    assert sorted({d.N for d in devs}) == [61, 113, 197, 317, 509] and len(devs) == 48
    m = ng.Multi(devs)
    T = m.transmission(Es)
    assert m.last_status == 0
    for wl, d, E, t in zip(wls, devs, Es, T):
        nx, ny = wl.basis()
        a = wl.atomic()
        kin = O.kinetic(nx, ny, a.LX, a.LY, 0.0, 0.0)
        n_open = np.array([np.sum(2 * e - kin > 0) for e in E])
        assert np.all(t >= -1e-12) and np.all(t <= n_open + 1e-6), d.N
    picks = [i for i, wl in enumerate(wls) if wl.extra["N"] == 61][:1]
    picks += [min((i for i, wl in enumerate(wls) if wl.extra["N"] == 509), key=lambda i: wls[i].extra["N_z"])]
    for i in picks:
        wl = wls[i]
        NGX, NGY = wl.ngx()
        a = O.atomic_from_input(**{k: v for k, v in wl.inp.items()
                                   if k in ("LX", "LY", "LZ", "ENCUT", "ETA", "TRANSMISSION_GRID")})
        nx, ny = O.planewave_basis(a.ENCUT, a.LX, a.LY)
        s = O.System(a, nx, ny, vrecs[i], NGX, NGY)
        for j in (20, 51):
            ref = O.transmission_tau(s, Es[i][j], 0, faithful_solver=False)
            assert T[i][j] == pytest.approx(ref, rel=RTOL, abs=1e-12), (wl.extra, j)
    sl = [np.linspace(0, d.N_z - 1, 8).round().astype(int) for d in devs]
    E2 = [e[[30, 50]] for e in Es]
    L = m.ldos(E2, sl)
    assert m.last_status == 0
    for d, g in zip(devs, L):
        for q in range(g.shape[2]):
            for e in range(g.shape[3]):
                B = g[:, :, q, e]
                w = np.linalg.eigvalsh((B - B.conj().T) / 2j)
                assert w.min() >= -1e-9 * max(1.0, np.abs(w).max()), (d.N, q, e)
    for i in picks:
        ref = devs[i].ldos_matrix(E2[i])[:, :, sl[i], :]
        assert np.max(np.abs(L[i] - ref)) <= 1e-12 * np.abs(ref).max()
    m.close()
    for d in devs:
        d.close()
