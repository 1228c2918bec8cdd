"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Bar: relative 1e-10 for every FP64 output (the same tolerance is written in each
assertion), on seeded inputs at sizes the oracle finishes in seconds, plus the
full-size metric configuration (N = 317, N_z = 1000) on two energies.
"""
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


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# ------------------------------------------------------------------ a7, a8
@pytest.mark.parametrize("N", [1, 5, 16, 17, 61, 100, 317, 600, 797])
def test_inverse_batch(ng, N):
    rng = np.random.default_rng(N)
    A = rand_c(rng, N, N, 3)
    A[:, :, 1] = A[:, :, 1] * 1e-3 + np.eye(N)[::-1]      # forces row interchanges
    inv = ng.inverse(A)
    for b in range(3):
        ref = np.linalg.inv(A[:, :, b])
        assert rel_err(inv[:, :, b], ref) < RTOL * max(1.0, np.linalg.cond(A[:, :, b]) / 1e3)
        assert rel_err(inv[:, :, b], O.inverse(A[:, :, b])) < RTOL * max(1.0, np.linalg.cond(A[:, :, b]) / 1e3)


@pytest.mark.parametrize("N", [3, 61, 317])
def test_matrix_product_batch(ng, N):
    rng = np.random.default_rng(7 + N)
    A, B = rand_c(rng, N, N, 2), rand_c(rng, N, N, 2)
    C = ng.matrix_product(A, B)
    for b in range(2):
        assert rel_err(C[:, :, b], O.matrix_product(A[:, :, b], B[:, :, b])) < RTOL


def test_singular_block_reports_status(ng):
    A = np.zeros((4, 4, 1), dtype=np.complex128)
    import negftrans_amd._lib as L
    out = np.zeros_like(A, order="F")
    st = L.load().negf_inverse_batch(4, 1, L.dptr(np.asfortranarray(A)), L.dptr(out))
    assert st == 1


# ------------------------------------------------------------------ a9
@pytest.mark.parametrize("N,N_z,DFL", [(13, 7, "DFL"), (5, 2, "DFL"), (61, 12, "DF"), (29, 9, "L")])
def test_tri_solver(ng, N, N_z, DFL):
    rng = np.random.default_rng(N * 31 + N_z)
    D = rand_c(rng, N_z, N, N) * 0.5 + 3.0 * np.eye(N)
    ref = O.tri_solver(D, DFL)
    out = ng.greensfunction_tri_solver(np.moveaxis(D, 0, 2), DFL)
    assert out["status"] == 0
    for key in ("diagonal", "first_column", "last_column"):
        r = np.moveaxis(ref[key], 0, 2)
        if np.any(r != 0):
            assert rel_err(out[key], r) < RTOL, key
        else:
            assert np.all(out[key] == 0), key


# ------------------------------------------------------------------ small systems
def small_system(N_z=8, seed=3, R2=6.5, NK=(1, 1), eta=1e-6, VL=0.0, VR=0.0, full=False):
    L = 12.0
    ENCUT = R2 * (2 * np.pi / L) ** 2 / 2
    nx, ny = O.planewave_basis(ENCUT, L, L)
    NGX = NGY = 2 * int(np.ceil(np.sqrt(R2))) + 1
    N_x = N_y = 2 * NGX + 4
    V = 0.02 * np.random.default_rng(seed).standard_normal((N_x, N_y, N_z)) + 0.01
    Vrec = O.v_reciprocal_all(V, NGX, NGY, full_window=full)
    a = O.Atomic(V_L=VL, V_R=VR, ETA=eta, LX=L, LY=L, LZ=N_z * 0.7, ENCUT=ENCUT)
    kx, ky, kw = O.kpoint_mesh(NK[0], NK[1], L, L)
    return O.System(a, nx, ny, Vrec, NGX, NGY, kx, ky, kw), V


def device_for(ng, s):
    a = s.atomic
    at = ng.Atomic(V_L=a.V_L, V_R=a.V_R, ETA=a.ETA, LX=a.LX, LY=a.LY, LZ=a.LZ, ENCUT=a.ENCUT)
    return ng.NegfDevice(s.nx, s.ny, at, s.Vrec, s.NGX, s.NGY)


@pytest.mark.parametrize("NK,VL,VR,full", [((1, 1), 0.0, 0.0, False), ((2, 3), 0.0, 0.0, True),
                                           ((1, 1), 0.02, -0.01, True)])
def test_transmission_small(ng, NK, VL, VR, full):
    s, _ = small_system(N_z=9, NK=NK, VL=VL, VR=VR, full=full)
    E = np.linspace(0.05, 1.2, 17)
    ref = O.transmission(s, E)
    dev = device_for(ng, s)
    got = dev.transmission(E, (s.kx, s.ky, s.kw))
    assert dev.last_status == 0
    np.testing.assert_allclose(got, ref, rtol=RTOL, atol=1e-12 * max(1.0, np.abs(ref).max()))


def test_transmission_two_slices(ng):
    s, _ = small_system(N_z=2, seed=9)
    E = np.array([0.1, 0.5, 0.9])
    ref = O.transmission(s, E)
    got = device_for(ng, s).transmission(E)
    np.testing.assert_allclose(got, ref, rtol=RTOL, atol=1e-13)


# ------------------------------------------------------------------ a5, a6, a10 drop-ins
def test_gpu_blocks(ng):
    """Hamiltonian_construction, E_minus_H_construction and Transmission_solver drop-ins
    (context-free forms and the context forms) against the oracle's literal restatement."""
    import negftrans_amd._lib as L
    lib = L.load()
    s, _ = small_system(N_z=5, seed=11, NK=(2, 2), VL=0.01, VR=-0.02)
    a = s.atomic
    N, Nz = s.N, s.N_z
    nx = np.ascontiguousarray(s.nx, dtype=np.int32)
    ny = np.ascontiguousarray(s.ny, dtype=np.int32)
    vrec = np.asfortranarray(s.Vrec, dtype=np.complex128)
    dev = device_for(ng, s)
    for ik in range(len(s.kx)):
        kx, ky = float(s.kx[ik]), float(s.ky[ik])
        Href = O.hamiltonian(s.nx, s.ny, a.LX, a.LY, kx, ky, s.Vrec, s.NGX, s.NGY)   # (N_z, N, N)
        H = np.zeros((N, N, Nz), dtype=np.complex128, order="F")
        assert lib.negf_hamiltonian_direct(N, Nz, L.iptr(nx), L.iptr(ny), a.LX, a.LY, kx, ky, L.dptr(vrec),
                                           s.NGX, s.NGY, L.dptr(H)) == 0
        assert rel_err(H, np.moveaxis(Href, 0, 2)) < RTOL
        Hc = np.zeros_like(H)
        assert lib.negf_hamiltonian(dev._h, kx, ky, L.dptr(Hc)) == 0
        assert rel_err(Hc, np.moveaxis(Href, 0, 2)) < RTOL
        z = 0.37 + 1j * a.ETA
        Dref = O.e_minus_h(Href.copy(), z, s.nx, s.ny, a.LX, a.LY, a.LZ, kx, ky, a.V_L, a.V_R)
        D = np.zeros_like(H)
        assert lib.negf_e_minus_h_from_h(N, Nz, L.dptr(H), z.real, z.imag, L.iptr(nx), L.iptr(ny), a.LX, a.LY,
                                         a.LZ, kx, ky, a.V_L, a.V_R, L.dptr(D)) == 0
        assert rel_err(D, np.moveaxis(Dref, 0, 2)) < RTOL
        Dc = np.zeros_like(H)
        assert lib.negf_e_minus_h(dev._h, z.real, z.imag, kx, ky, L.dptr(Dc)) == 0
        assert rel_err(Dc, np.moveaxis(Dref, 0, 2)) < RTOL
        # Transmission_solver on the rescaled first column of the oracle's F sweep
        G = O.rgf_first_column_last_slice(Dref) * 2.0 * (a.LZ / Nz) ** 2
        tref = O.transmission_solver(G, z.real, s.nx, s.ny, a.LX, a.LY, a.LZ, Nz, kx, ky, a.V_L, a.V_R)
        Gf = np.asfortranarray(G)
        tau = np.zeros(1)
        assert lib.negf_transmission_trace(N, Nz, L.dptr(Gf), z.real, L.iptr(nx), L.iptr(ny), a.LX, a.LY,
                                           a.LZ, kx, ky, a.V_L, a.V_R, L.dptr(tau)) == 0
        assert tau[0] == pytest.approx(tref, rel=RTOL, abs=1e-14)


# ------------------------------------------------------------------ workloads
def oracle_system(wl, vrec, ngx, ngy):
    a = O.atomic_from_input(**{k: v for k, v in wl.inp.items()
                               if k in ("V_L", "V_R", "MU", "ETA", "TEMPERATURE", "LX", "LY", "LZ",
                                        "ENCUT", "GAP", "VDS", "TRANSMISSION_GRID")})
    nx, ny = O.planewave_basis(a.ENCUT, a.LX, a.LY)
    return O.System(a, nx, ny, vrec, ngx, ngy)


def test_vhat_c0(ng):
    from negftrans_amd import workloads
    wl = workloads.qpc_c0(n_energy=4)
    NGX, NGY = wl.ngx()
    V = wl.V_real_eV / O.HARTREE
    for full in (False, True):
        got = ng.local_potential_real_to_planewave(V, NGX, NGY, full_window=full)
        ref = O.v_reciprocal_all(V, NGX, NGY, full_window=full)
        assert rel_err(got, ref) < RTOL
        assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max()


def test_transmission_qpc_c0(ng):
    """Config 0 geometry (N = 61, N_z = 200), 24 of its 400 energies."""
    from negftrans_amd import workloads
    wl = workloads.qpc_c0()
    NGX, NGY = wl.ngx()
    vrec = ng.local_potential_real_to_planewave(wl.V_real_eV / O.HARTREE, NGX, NGY)
    s = oracle_system(wl, np.asarray(vrec), NGX, NGY)
    assert s.N == 61 and s.N_z == 200
    E = wl.energies_hartree()[::17][:24]
    ref = O.transmission(s, E)
    dev = ng.NegfDevice(s.nx, s.ny, wl.atomic(), vrec, NGX, NGY)
    got = dev.transmission(E)
    np.testing.assert_allclose(got, ref, rtol=RTOL, atol=1e-12)
    assert np.all(got >= 0)


def test_transmission_disorder_c1_full_size(ng):
    """Metric configuration at full size (N = 317, N_z = 1000): two energies against the
    oracle's O(N^2)-state restatement (same generators), plus size-independent bounds
    on 64 energies: 0 <= T <= number of open channels."""
    from negftrans_amd import workloads
    wl = workloads.disorder_c1()
    NGX, NGY = wl.ngx()
    vrec = ng.local_potential_real_to_planewave(wl.V_real_eV / O.HARTREE, NGX, NGY)
    a = wl.atomic()
    dev = ng.NegfDevice(*wl.basis(), a, vrec, NGX, NGY)
    assert dev.N == 317 and dev.N_z == 1000
    Eall = wl.energies_hartree()
    E = Eall[::64]
    got = dev.transmission(E)
    nx, ny = wl.basis()
    kin = O.kinetic(nx, ny, a.LX, a.LY, 0.0, 0.0)
    n_open = np.array([np.sum(2 * e - kin > 0) for e in E])
    assert np.all(got >= -1e-12) and np.all(got <= n_open + 1e-6)
    s = oracle_system(wl, np.asarray(vrec), NGX, NGY)
    for i in (13, 50):
        ref = O.transmission_tau(s, E[i], 0, faithful_solver=False)
        assert got[i] == pytest.approx(ref, rel=RTOL, abs=1e-12)


def test_transmission_one_sided_padded_320(ng, monkeypatch):
    """The form the benchmark times: one workgroup per energy (k_transmission), whose
    open-row product runs on the matrix cores at padded sizes 256-320.  N = 317 (padded
    320) on a short wire (N_z = 6) against the oracle at relative 1e-10, energies with up
    to ~70 open channels."""
    s, _ = small_system(N_z=6, seed=11, R2=100.5)
    assert len(s.nx) == 317
    E = np.array([0.3, 1.1, 2.7])
    ref = O.transmission(s, E)
    monkeypatch.setenv("NEGF_DIAG_ONE_SIDED", "1")
    dev = device_for(ng, s)
    monkeypatch.delenv("NEGF_DIAG_ONE_SIDED")
    got = dev.transmission(E)
    assert dev.last_status == 0
    np.testing.assert_allclose(got, ref, rtol=RTOL, atol=1e-12)
    assert ref.max() > 10.0


def test_transmission_disorder_c1_one_sided(ng, monkeypatch):
    """The benchmark's kernel at the metric configuration (N = 317, N_z = 1000), forced to
    one workgroup per energy: two energies against the oracle's O(N^2)-state
    restatement (same generators) at relative 1e-10."""
    from negftrans_amd import workloads
    wl = workloads.disorder_c1()
    NGX, NGY = wl.ngx()
    vrec = ng.local_potential_real_to_planewave(wl.V_real_eV / O.HARTREE, NGX, NGY)
    monkeypatch.setenv("NEGF_DIAG_ONE_SIDED", "1")
    dev = ng.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY)
    monkeypatch.delenv("NEGF_DIAG_ONE_SIDED")
    E = wl.energies_hartree()[[1234, 3777]]
    got = dev.transmission(E)
    assert dev.last_status == 0
    s = oracle_system(wl, np.asarray(vrec), NGX, NGY)
    for i in range(2):
        ref = O.transmission_tau(s, E[i], 0, faithful_solver=False)
        assert got[i] == pytest.approx(ref, rel=RTOL, abs=1e-12)


def test_transmission_c3_full_depth(ng, monkeypatch):
    """The largest configuration at its own size (N = 797, padded 800: the four-tile
    matrix-core Gauss-Jordan; N_z = 2000), where the oracle would take minutes per
    energy: two energies (5 and 9 open channels) through the two-sided sweep and, forced,
    the one-sided sweep.  These are separate code paths (two workgroups meeting at the
    middle slice, against one workgroup over all 2000 slices), so they must agree to
    relative 1e-9, and 0 <= T <= n_open."""
    from negftrans_amd import workloads
    wl = workloads.disorder_c3()
    NGX, NGY = wl.ngx()
    vrec = ng.local_potential_real_to_planewave(wl.V_real_eV / O.HARTREE, NGX, NGY)
    nx, ny = wl.basis()
    a = wl.atomic()
    E = wl.energies_hartree()[[600, 1023]]
    kin = O.kinetic(np.asarray(nx), np.asarray(ny), a.LX, a.LY, 0.0, 0.0)
    n_open = np.array([np.sum(2 * e - kin > 0) for e in E])
    assert list(n_open) == [5, 9]
    dev = ng.NegfDevice(nx, ny, a, vrec, NGX, NGY)
    assert (dev.N, dev.N_z) == (797, 2000)
    two = dev.transmission(E)
    assert dev.last_status == 0
    assert dev.last_variant == 3                  # two-sided, four-tile
    dev.close()
    monkeypatch.setenv("NEGF_DIAG_ONE_SIDED", "1")
    dev = ng.NegfDevice(nx, ny, a, vrec, NGX, NGY)
    monkeypatch.delenv("NEGF_DIAG_ONE_SIDED")
    one = dev.transmission(E)
    assert dev.last_status == 0
    assert dev.last_variant == 2                  # one-sided, four-tile
    dev.close()
    np.testing.assert_allclose(two, one, rtol=1e-9, atol=1e-12)
    assert np.all(two >= -1e-9) and np.all(two <= n_open + 1e-9)


def test_density_c1_full_size(ng):
    """D mode at the metric configuration's size (N = 317, N_z = 1000): one contour node
    of the density (z in the upper half plane, weight 1) against the oracle's D-mode
    restatement on every slice, relative 1e-10.  The kernel keeps all 1000 generators
    (1.64 GB of workspace for the job)."""
    from negftrans_amd import workloads
    wl = workloads.disorder_c1()
    NGX, NGY = wl.ngx()
    vrec = ng.local_potential_real_to_planewave(wl.V_real_eV / O.HARTREE, NGX, NGY)
    dev = ng.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY)
    z = wl.mstar * complex(0.015, 0.004) / O.HARTREE
    got = dev.density_matrix(np.array([z]), np.array([1.0 + 0j]))
    assert dev.last_status == 0
    ref = O.diagonal_blocks(oracle_system(wl, np.asarray(vrec), NGX, NGY), z, 0)
    assert rel_err(got, np.moveaxis(ref, 0, 2)) < RTOL
    # a second call gives the same bits (two workgroups per job, no cross-run race)
    assert np.array_equal(dev.density_matrix(np.array([z]), np.array([1.0 + 0j])), got)


# ------------------------------------------------------------------ a12, a13, a14
def test_density_small(ng):
    s, V = small_system(N_z=5, seed=21, NK=(2, 1), eta=1e-4)
    a = s.atomic
    a.MU, a.TEMPERATURE = 0.6, 300.0
    a.GAP = 6 * O.K_B * a.TEMPERATURE
    minV = float(V.min())
    ref = O.density_matrix(s, 7, 5, minV)
    at = ng.Atomic(V_L=a.V_L, V_R=a.V_R, ETA=a.ETA, LX=a.LX, LY=a.LY, LZ=a.LZ, ENCUT=a.ENCUT,
                   MU=a.MU, TEMPERATURE=a.TEMPERATURE, GAP=a.GAP)
    z, w = ng.density_contour(at, 7, 5, minV)
    dev = device_for(ng, s)
    got = dev.density_matrix(z, w, (s.kx, s.ky, s.kw))
    assert rel_err(got, np.moveaxis(ref, 0, 2)) < RTOL


@pytest.mark.parametrize("N_z", [2, 3, 7])
def test_diagonal_two_sided_and_one_sided(ng, N_z, monkeypatch):
    """D mode: a batch of few jobs runs two workgroups per job (generators from both
    leads, diagonal blocks outwards from the cut at N_z / 2); NEGF_DIAG_ONE_SIDED=1 keeps
    the one-workgroup backward-generator form.  Both against the oracle's tri_solver 'D'
    restatement, relative 1e-10, down to N_z = 2 (one slice per side; contexts need
    N_z >= 2)."""
    s, _ = small_system(N_z=N_z, seed=31, eta=1e-4)
    zs = np.array([0.3 + 0.05j, 0.9 + 0.002j, 1.4 + 0.3j])
    dev = device_for(ng, s)
    two = dev.density_matrix(zs, np.ones(3, dtype=complex))
    assert dev.last_status == 0
    monkeypatch.setenv("NEGF_DIAG_ONE_SIDED", "1")
    dev1 = device_for(ng, s)
    monkeypatch.delenv("NEGF_DIAG_ONE_SIDED")
    one = dev1.density_matrix(zs, np.ones(3, dtype=complex))
    assert dev1.last_status == 0
    ref = sum(O.diagonal_blocks(s, z, 0) for z in zs)
    assert rel_err(two, np.moveaxis(ref, 0, 2)) < RTOL
    assert rel_err(one, np.moveaxis(ref, 0, 2)) < RTOL
    assert rel_err(two, one) < RTOL


@pytest.mark.parametrize("N_z,NK", [(2, (1, 1)), (3, (2, 1)), (8, (1, 1)), (11, (1, 2))])
def test_transmission_two_sided_and_one_sided(ng, N_z, NK, monkeypatch):
    """Transmission: a call with few jobs runs two workgroups per job -- forward
    generators and the open-column chain of G_{k,1} from the left lead, backward
    generators and the open-row chain from the right, joined at the cut N_z / 2 through
    (I - f_m g_{m+1})^{-1}; NEGF_DIAG_ONE_SIDED=1 keeps the one-workgroup backward sweep.
    Both against the oracle (relative 1e-10), and against each other, down to N_z = 2;
    energies with no open channel give 0 exactly in both."""
    s, _ = small_system(N_z=N_z, seed=41, NK=NK, eta=1e-4)
    E = np.concatenate([[-0.2], np.linspace(0.05, 1.2, 9)])
    kp = (s.kx, s.ky, s.kw)
    ref = O.transmission(s, E)
    dev = device_for(ng, s)
    two = dev.transmission(E, kp)
    assert dev.last_status == 0
    monkeypatch.setenv("NEGF_DIAG_ONE_SIDED", "1")
    dev1 = device_for(ng, s)
    monkeypatch.delenv("NEGF_DIAG_ONE_SIDED")
    one = dev1.transmission(E, kp)
    assert dev1.last_status == 0
    tol = dict(rtol=RTOL, atol=1e-12 * max(1.0, np.abs(ref).max()))
    np.testing.assert_allclose(two, ref, **tol)
    np.testing.assert_allclose(one, ref, **tol)
    np.testing.assert_allclose(two, one, **tol)
    assert two[0] == 0.0 and one[0] == 0.0 and np.abs(ref).max() > 0.5


def test_ldos_small(ng):
    s, V = small_system(N_z=6, seed=22, NK=(1, 2), eta=1e-3)
    E = np.array([0.2, 0.7])
    dev = device_for(ng, s)
    got = dev.ldos_matrix(E, (s.kx, s.ky, s.kw))
    for i, e in enumerate(E):
        ref = O.ldos_matrix(s, e)
        assert rel_err(got[:, :, :, i], np.moveaxis(ref, 0, 2)) < RTOL


def test_neq_density_small(ng):
    """Bias-window part of the non-equilibrium density (extension) against the oracle's
    dense-inversion restatement, two k points, Simpson and Gauss-Legendre windows."""
    s, V = small_system(N_z=6, seed=23, NK=(2, 1), eta=1e-4, VL=0.01, VR=0.0)
    a = s.atomic
    a.MU, a.TEMPERATURE = 0.5, 300.0
    a.GAP = 6 * O.K_B * a.TEMPERATURE
    dev = device_for(ng, s)
    ha = ng.Atomic(MU=a.MU, TEMPERATURE=a.TEMPERATURE, GAP=a.GAP)
    for quad in ("simpson", "gauss-legendre"):
        E, w = ng.bias_window(ha, 0.02, 9, quad)
        got = dev.neq_density_matrix(E, w, (s.kx, s.ky, s.kw))
        assert dev.last_status == 0
        ref = np.zeros((s.N_z, s.N, s.N), dtype=np.complex128)
        for ik in range(len(s.kx)):
            for e, q in zip(E, w):
                ref += s.kw[ik] * q * O.neq_spectral_blocks(s, float(e), ik, "L")
        assert rel_err(got, np.moveaxis(ref, 0, 2)) < RTOL


@pytest.mark.parametrize("N_z", [2, 3, 7])
def test_neq_two_sided_and_one_sided(ng, N_z, monkeypatch):
    """Bias window with two workgroups per job (forward generators and the left parts'
    first columns on one side, backward generators on the other, outwards from the cut
    at N_z / 2) and with NEGF_DIAG_ONE_SIDED=1 (k_neq), both against the oracle's
    dense-inversion restatement, relative 1e-10, down to N_z = 2."""
    s, _ = small_system(N_z=N_z, seed=41, eta=1e-4, VL=0.01, VR=0.0)
    a = s.atomic
    a.MU, a.TEMPERATURE = 0.5, 300.0
    a.GAP = 6 * O.K_B * a.TEMPERATURE
    ha = ng.Atomic(MU=a.MU, TEMPERATURE=a.TEMPERATURE, GAP=a.GAP)
    E, w = ng.bias_window(ha, 0.02, 5, "gauss-legendre")
    dev = device_for(ng, s)
    two = dev.neq_density_matrix(E, w)
    assert dev.last_status == 0
    monkeypatch.setenv("NEGF_DIAG_ONE_SIDED", "1")
    dev1 = device_for(ng, s)
    monkeypatch.delenv("NEGF_DIAG_ONE_SIDED")
    one = dev1.neq_density_matrix(E, w)
    assert dev1.last_status == 0
    ref = np.zeros((s.N_z, s.N, s.N), dtype=np.complex128)
    for e, q in zip(E, w):
        ref += q * O.neq_spectral_blocks(s, float(e), 0, "L")
    assert np.abs(ref).max() > 0
    assert rel_err(two, np.moveaxis(ref, 0, 2)) < RTOL
    assert rel_err(one, np.moveaxis(ref, 0, 2)) < RTOL
    assert rel_err(two, one) < RTOL


def test_results_bit_reproducible(ng):
    """Every path returns bit-identical results when called twice on the same inputs:
    transmission (per-job results summed on the host), and the two-workgroup D mode
    (density, LDOS) and bias window with one job per call.  Within one job every output
    element is written by one workgroup, so any difference would point at a race.  (With
    several jobs the FP64 atomic accumulation into shared output blocks is order-dependent
    in the last bits, which is expected.)"""
    s, _ = small_system(N_z=9, seed=51, eta=1e-4, VL=0.01, VR=0.0)
    a = s.atomic
    a.MU, a.TEMPERATURE = 0.5, 300.0
    a.GAP = 6 * O.K_B * a.TEMPERATURE
    dev = device_for(ng, s)
    E = np.linspace(0.05, 1.2, 9)
    zs = np.array([0.3 + 0.05j, 0.9 + 0.002j])
    ha = ng.Atomic(MU=a.MU, TEMPERATURE=a.TEMPERATURE, GAP=a.GAP)
    Eb, wb = ng.bias_window(ha, 0.02, 5, "gauss-legendre")
    mid = len(Eb) // 2
    runs = []
    for _ in range(2):
        runs.append((dev.transmission(E), dev.density_matrix(zs[1:], np.ones(1, dtype=complex)),
                     dev.ldos_matrix(E[4:5]), dev.neq_density_matrix(Eb[mid:mid + 1], wb[mid:mid + 1])))
    assert np.abs(runs[0][3]).max() > 0
    for x, y in zip(*runs):
        assert np.array_equal(x, y)


def test_pw_to_real(ng):
    L = 6.0
    nx, ny = O.planewave_basis(4.5 * (2 * np.pi / L) ** 2 / 2, L, L)
    N = len(nx)
    rng = np.random.default_rng(5)
    M = rand_c(rng, N, N, 3)
    got = ng.greensfunction_planewave_to_real(M, nx, ny, 6, 8)
    for b in range(3):
        ref = O.pw_to_real_literal(M[:, :, b], nx, ny, 6, 8)
        assert rel_err(got[:, :, b], ref) < RTOL
    nx, ny = O.planewave_basis(100.5 * (2 * np.pi / L) ** 2 / 2, L, L)
    M = rand_c(rng, len(nx), len(nx), 2)
    got = ng.greensfunction_planewave_to_real(M, nx, ny, 64, 64)
    for b in range(2):
        assert rel_err(got[:, :, b], O.pw_to_real(M[:, :, b], nx, ny, 64, 64)) < RTOL


# ------------------------------------------------------------------ f4: unused reference path
@pytest.mark.parametrize("Nx,Ny,Rx,Ry", [(6, 4, 2, 1), (16, 12, 5, 4), (8, 8, 4, 4), (32, 32, 8, 8)])
def test_nonlocal_potential(ng, Nx, Ny, Rx, Ry):
    """NonLocalPotential_RealToPlanewave (src/grid.f90:137-234) on the GPU against the
    oracle (its FFT restatement, pinned by the literal four-fold sum and the local-potential
    identity), including the |f| = N/2 edge (periodic value)."""
    rng = np.random.default_rng(Nx * 7 + Ny)
    V = rand_c(rng, Nx, Ny, Nx, Ny)
    got = ng.nonlocal_potential_real_to_planewave(V, Rx, Ry)
    ref = O.nonlocal_potential_r2p(V, Rx, Ry)
    assert rel_err(got, ref) < RTOL


@pytest.mark.parametrize("R2", [191.5, 318.5])
def test_transmission_wide_blocks(ng, R2):
    """Blocks above Np = 512 run the four-tile, 16-column matrix-core Gauss-Jordan:
    N = 593 (padded 608: the last wave owns two of its four tiles) and N = 997 (padded
    1024: sixteen full waves), three slices, against the oracle, relative 1e-10."""
    s, _ = small_system(N_z=3, seed=61, R2=R2, eta=1e-4)
    assert s.N in (593, 997)
    kin = O.kinetic(s.nx, s.ny, s.atomic.LX, s.atomic.LY, 0.0, 0.0)
    E = np.array([0.3, 0.9]) * kin.max() * 0.05
    ref = O.transmission(s, E)
    got = device_for(ng, s).transmission(E)
    assert np.abs(ref).max() > 0
    np.testing.assert_allclose(got, ref, rtol=RTOL, atol=1e-12 * max(1.0, np.abs(ref).max()))


def test_wide_blocks_d_mode_and_bias_window(ng, monkeypatch):
    """D mode (two-sided and one-sided) and the bias window above Np = 512 (N = 593,
    padded 608, four-tile matrix-core inversion), three slices, against the oracle,
    relative 1e-10."""
    s, _ = small_system(N_z=3, seed=62, R2=191.5, eta=1e-4, VL=0.01, VR=0.0)
    assert s.N == 593
    kin = O.kinetic(s.nx, s.ny, s.atomic.LX, s.atomic.LY, 0.0, 0.0)
    z = complex(0.4 * 0.05 * kin.max(), 0.01)
    ref = np.moveaxis(O.diagonal_blocks(s, z, 0), 0, 2)
    dev = device_for(ng, s)
    assert rel_err(dev.density_matrix(np.array([z]), np.ones(1, dtype=complex)), ref) < RTOL
    monkeypatch.setenv("NEGF_DIAG_ONE_SIDED", "1")
    dev1 = device_for(ng, s)
    monkeypatch.delenv("NEGF_DIAG_ONE_SIDED")
    assert rel_err(dev1.density_matrix(np.array([z]), np.ones(1, dtype=complex)), ref) < RTOL
    e = 0.4 * 0.05 * kin.max()
    refb = np.moveaxis(O.neq_spectral_blocks(s, e, 0, "L"), 0, 2)
    assert np.abs(refb).max() > 0
    got = dev.neq_density_matrix(np.array([e]), np.ones(1, dtype=complex))
    assert rel_err(got, refb) < RTOL


def test_transmission_c3_block_size(ng):
    """The largest configuration's block size (N = 797, padded 800: four-tile waves,
    16-column panels) on the disorder of that configuration: against the oracle on the
    first 6 slices (2 energies, relative 1e-10), and at N_z = 40 the mirror property --
    with V_L = V_R, reversing the potential in z leaves T unchanged (two-terminal
    reciprocity; the tolerance allows the O(eta) loss of unitarity) -- with
    0 <= T <= open channels."""
    from negftrans_amd import workloads
    wl = workloads.disorder_c3(N_z=40)
    NGX, NGY = wl.ngx()
    vrec = np.asfortranarray(ng.local_potential_real_to_planewave(wl.V_real_eV / O.HARTREE, NGX, NGY))
    a = wl.atomic()
    nx, ny = wl.basis()
    assert len(nx) == 797
    E = wl.energies_hartree()[[700, 1000]]       # 5 and 9 open channels
    short = np.asfortranarray(vrec[:, :, :6])
    a6 = ng.Atomic(V_L=a.V_L, V_R=a.V_R, ETA=a.ETA, LX=a.LX, LY=a.LY, LZ=a.LZ * 6 / 40, ENCUT=a.ENCUT)
    got6 = ng.NegfDevice(nx, ny, a6, short, NGX, NGY).transmission(E)
    s6 = O.System(O.Atomic(V_L=a.V_L, V_R=a.V_R, ETA=a.ETA, LX=a.LX, LY=a.LY, LZ=a.LZ * 6 / 40, ENCUT=a.ENCUT),
                  nx, ny, np.asarray(short), NGX, NGY)
    np.testing.assert_allclose(got6, O.transmission(s6, E), rtol=RTOL, atol=1e-12)
    fwd = ng.NegfDevice(nx, ny, a, vrec, NGX, NGY).transmission(E)
    rev = ng.NegfDevice(nx, ny, a, np.asfortranarray(vrec[:, :, ::-1]), NGX, NGY).transmission(E)
    kin = O.kinetic(nx, ny, a.LX, a.LY, 0.0, 0.0)
    n_open = np.array([np.sum(2 * e - kin > 0) for e in E])
    assert np.all(fwd >= -1e-12) and np.all(fwd <= n_open + 1e-6) and n_open.min() >= 2
    np.testing.assert_allclose(rev, fwd, rtol=1e-6, atol=1e-9)


def test_empty_and_closed_batches(ng):
    """Edge cases of the batched calls: no energies, no contour nodes, and energies below
    every channel (no open channel: T = 0 exactly, as the reference's sum of zeros)."""
    s, _ = small_system(N_z=5, seed=31)
    dev = device_for(ng, s)
    assert dev.transmission(np.array([])).shape == (0,)
    dm = dev.density_matrix(np.array([], dtype=np.complex128), np.array([], dtype=np.complex128))
    assert np.all(dm == 0)
    assert dev.ldos_matrix(np.array([])).shape[-1] == 0
    T = dev.transmission(np.array([-0.3, -0.01, 0.0]))
    assert np.all(T == 0.0)
    # an uneven batch with a device that has no energies
    s2, _ = small_system(N_z=7, seed=32)
    d2 = device_for(ng, s2)
    E2 = np.linspace(0.1, 0.9, 5)
    got = ng.Multi([dev, d2]).transmission([np.array([]), E2])
    assert got[0].shape == (0,)
    np.testing.assert_allclose(got[1], d2.transmission(E2), rtol=1e-12)
