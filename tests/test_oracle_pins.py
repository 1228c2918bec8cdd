"""Pins of the CPU oracle against brute force, closed forms and invariants.

The reference ships no fixtures or golden outputs, and it cannot be built here
(gfortran, FFTW3 and a LAPACK with the standard symbol names are absent), so
every oracle function is checked against an independent computation chosen so
that a dropped term, a wrong sign or a transposed operand fails.
"""
import math

import numpy as np
import pytest

from oracle import negf_oracle as O


def rng(seed=0):
    return np.random.default_rng(seed)


# ------------------------------------------------------------------ basis / mesh
@pytest.mark.parametrize("R2,expected", [(19, 61), (36.5, 113), (64.5, 197), (100.5, 317),
                                         (162.5, 509), (256.5, 797)])
def test_basis_gauss_circle_counts(R2, expected):
    L = 400.0 / O.A_0
    ENCUT = R2 * (2 * O.PI / L) ** 2 / 2
    nx, ny = O.planewave_basis(ENCUT, L, L)
    # independent count: integer points with i^2 + j^2 <= R2
    g = np.arange(-30, 31)
    I, J = np.meshgrid(g, g, indexing="ij")
    assert len(nx) == expected == int(np.sum(I ** 2 + J ** 2 <= R2 + 1e-9))
    # ordering rule: j outer ascending, i inner ascending (src/grid.f90:420-428)
    key = ny.astype(np.int64) * 1000 + nx
    assert np.all(np.diff(key) > 0)


def test_basis_rectangular_cell():
    Lx, Ly = 8.0 / O.A_0, 18.0 / O.A_0
    ENCUT = 100.0 / O.HARTREE
    nx, ny = O.planewave_basis(ENCUT, Lx, Ly)
    e = ((2 * np.pi * nx / Lx) ** 2 + (2 * np.pi * ny / Ly) ** 2) / 2
    assert np.all(e <= ENCUT)
    # every lattice point inside the cutoff is present
    cnt = 0
    for j in range(-20, 21):
        for i in range(-20, 21):
            if ((2 * np.pi * i / Lx) ** 2 + (2 * np.pi * j / Ly) ** 2) / 2 <= ENCUT:
                cnt += 1
    assert cnt == len(nx)


def test_kpoint_mesh_monkhorst_pack():
    L = 10.0
    kx, ky, w = O.kpoint_mesh(3, 2, L, L)
    assert np.isclose(w.sum(), 1.0)
    np.testing.assert_allclose(sorted(set(np.round(kx, 12))),
                               np.array([-1, 0, 1]) / 3 * 2 * np.pi / L, atol=1e-12)
    np.testing.assert_allclose(sorted(set(np.round(ky, 12))),
                               np.array([-0.5, 0.5]) / 2 * 2 * np.pi / L, atol=1e-12)
    # x fastest (index = (j-1)*NKX + i)
    assert kx[0] < kx[1] < kx[2] and ky[0] == ky[2] < ky[3]


def test_ngx_auto_rule_uses_LY_in_the_test():
    # LY large so the NGX test (against LY) triggers even though NGX input is big
    NGX, NGY = O.ngx_auto(12, 0, 8.0, 18.0, 100.0)
    sx = math.ceil((8.0 / O.A_0) / O.PI * math.sqrt(100.0 / O.HARTREE / 2) * 2)
    sy = math.ceil((18.0 / O.A_0) / O.PI * math.sqrt(100.0 / O.HARTREE / 2) * 2)
    assert sy >= 12 and NGX == sx + 1 and NGY == sy + 1


# ------------------------------------------------------------------ stage 1: V-hat
def test_vhat_constant_and_cosine():
    N_x, N_y = 16, 12
    V = np.full((N_x, N_y), 0.37)
    out = O.local_potential_r2p(V, 4, 3)
    ref = np.zeros_like(out)
    ref[4, 3] = 0.37
    np.testing.assert_allclose(out, ref, atol=1e-15)
    x = np.arange(N_x)[:, None]
    y = np.arange(N_y)[None, :]
    V = np.cos(2 * np.pi * 2 * x / N_x) + 0 * y + 0.5 * np.sin(2 * np.pi * 1 * y / N_y)
    out = O.local_potential_r2p(V, 4, 3)
    assert np.isclose(out[4 + 2, 3], 0.5) and np.isclose(out[4 - 2, 3], 0.5)
    # sin(2 pi y/N_y) = (e^{+i} - e^{-i})/(2i): coefficient at q=+1 is 1/(2i) * 0.5
    assert np.isclose(out[4, 3 + 1], 0.5 / 2j) and np.isclose(out[4, 3 - 1], -0.5 / 2j)
    others = np.abs(out).copy()
    for idx in [(6, 3), (2, 3), (4, 4), (4, 2)]:
        others[idx] = 0
    assert others.max() < 1e-14


def test_vhat_literal_dft_and_hermiticity():
    N_x, N_y, NGX, NGY = 10, 8, 4, 3
    V = rng(1).standard_normal((N_x, N_y))
    out = O.local_potential_r2p(V, NGX, NGY)
    for p in range(-NGX, NGX + 1):
        for q in range(-NGY, NGY + 1):
            s = sum(V[x, y] * np.exp(-2j * np.pi * (p * x / N_x + q * y / N_y))
                    for x in range(N_x) for y in range(N_y)) / (N_x * N_y)
            assert abs(out[p + NGX, q + NGY] - s) < 1e-13
            assert abs(out[p + NGX, q + NGY] - np.conj(out[-p + NGX, -q + NGY])) < 1e-14


def test_vhat_window_truncation():
    N_x, N_y, N_z, NGX, NGY = 12, 12, 2, 5, 5
    V = rng(2).standard_normal((N_x, N_y, N_z))
    full = O.v_reciprocal_all(V, NGX, NGY, full_window=True)
    half = O.v_reciprocal_all(V, NGX, NGY)
    w = np.zeros_like(full, dtype=bool)
    w[NGX - 2:NGX + 3, NGY - 2:NGY + 3, :] = True   # |p|,|q| <= NGX/2 = 2
    np.testing.assert_array_equal(half[w], full[w])
    assert np.all(half[~w] == 0) and np.any(full[~w] != 0)


# ------------------------------------------------------------------ blocks
def tiny_system(N_z=5, seed=3, R2=2.5, V_L=0.0, V_R=0.0, NK=(1, 1), eta=1e-3):
    L = 6.0
    ENCUT = R2 * (2 * np.pi / L) ** 2 / 2
    nx, ny = O.planewave_basis(ENCUT, L, L)
    NGX = NGY = 2 * int(math.ceil(math.sqrt(R2))) + 1
    N_x = N_y = 2 * NGX + 2
    V = 0.05 * rng(seed).standard_normal((N_x, N_y, N_z))
    Vrec = O.v_reciprocal_all(V, NGX, NGY, full_window=True)
    a = O.Atomic(V_L=V_L, V_R=V_R, ETA=eta, LX=L, LY=L, LZ=N_z * 0.9, ENCUT=ENCUT)
    kx, ky, kw = O.kpoint_mesh(NK[0], NK[1], L, L)
    return O.System(a, nx, ny, Vrec, NGX, NGY, kx, ky, kw), V


def test_hamiltonian_and_blocks_literal():
    sys_, _ = tiny_system()
    a = sys_.atomic
    kx, ky = 0.07, -0.03
    H = O.hamiltonian(sys_.nx, sys_.ny, a.LX, a.LY, kx, ky, sys_.Vrec, sys_.NGX, sys_.NGY)
    N, N_z = sys_.N, sys_.N_z
    for k in range(N_z):
        for i in range(N):
            for j in range(N):
                v = sys_.Vrec[sys_.nx[i] - sys_.nx[j] + sys_.NGX, sys_.ny[i] - sys_.ny[j] + sys_.NGY, k]
                if i == j:
                    v += ((2 * sys_.nx[i] * np.pi / a.LX + kx) ** 2
                          + (2 * sys_.ny[i] * np.pi / a.LY + ky) ** 2) / 2
                assert H[k, i, j] == pytest.approx(v, abs=1e-15)
    E = 0.31 + 0.002j
    D = O.e_minus_h(H, E, sys_.nx, sys_.ny, a.LX, a.LY, a.LZ, kx, ky, 0.01, 0.02)
    d = a.LZ / N_z
    I = np.eye(N)
    for k in range(N_z):
        ref = 2 * d * d * (E * I - H[k]) - 2 * I
        if k == 0:
            ref = ref + np.diag(np.exp(1j * d * np.sqrt(2 * (E - 0.01) - 2 * np.diag(H[k]).real
                                                         + 2 * sys_.Vrec[sys_.NGX, sys_.NGY, k].real + 0j)))
        if k == N_z - 1:
            ref = ref + np.diag(np.exp(1j * d * np.sqrt(2 * (E - 0.02) - 2 * np.diag(H[k]).real
                                                         + 2 * sys_.Vrec[sys_.NGX, sys_.NGY, k].real + 0j)))
        np.testing.assert_allclose(D[k], ref, rtol=1e-12, atol=1e-14)


# ------------------------------------------------------------------ stage 3
def test_inverse_and_product():
    A = rng(4).standard_normal((23, 23)) + 1j * rng(5).standard_normal((23, 23))
    B = rng(6).standard_normal((23, 23)) + 1j * rng(7).standard_normal((23, 23))
    np.testing.assert_allclose(O.inverse(A) @ A, np.eye(23), atol=1e-11)
    np.testing.assert_allclose(O.matrix_product(A, B), A @ B, rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("N,N_z", [(5, 2), (5, 3), (9, 7), (13, 12)])
def test_tri_solver_vs_dense_inverse(N, N_z):
    r = rng(N * 100 + N_z)
    D = (r.standard_normal((N_z, N, N)) + 1j * r.standard_normal((N_z, N, N))) * 0.6
    D += 3.0 * np.eye(N)
    full = np.linalg.inv(O.assemble_dense(D))
    out = O.tri_solver(D, "DFL")
    for k in range(N_z):
        blk = lambda r_, c_: full[r_ * N:(r_ + 1) * N, c_ * N:(c_ + 1) * N]
        np.testing.assert_allclose(out["diagonal"][k], blk(k, k), rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(out["first_column"][k], blk(k, 0), rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(out["last_column"][k], blk(k, N_z - 1), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(O.rgf_first_column_last_slice(D), full[(N_z - 1) * N:, :N],
                               rtol=1e-10, atol=1e-12)
    only_f = O.tri_solver(D, "F")
    assert np.all(only_f["diagonal"] == 0) and np.all(only_f["last_column"] == 0)


# ------------------------------------------------------------------ stage 4: trace
def chain_T_closed_form(E, eta, kin, delta, N_z, V_L=0.0, V_R=0.0):
    """1-D chain with unit hopping: G_{n1} = (-1)^{n+1}/det(A), det multilinear in the
    two end corrections; D_n by the three-term recurrence.  Independent of LAPACK."""
    z = E + 1j * eta
    d = 2 * delta ** 2 * (z - kin / 2) - 2
    sL = np.exp(1j * np.sqrt(2 * (z - V_L) - kin + 0j) * delta)
    sR = np.exp(1j * np.sqrt(2 * (z - V_R) - kin + 0j) * delta)
    Dm = [1.0 + 0j, d]
    for _ in range(2, N_z + 1):
        Dm.append(d * Dm[-1] - Dm[-2])
    det = Dm[N_z] + (sL + sR) * Dm[N_z - 1] + sL * sR * Dm[N_z - 2]
    G = (-1) ** (N_z + 1) / det * 2 * delta ** 2
    gL = np.sin(np.sqrt(max(2 * (E - V_L) - kin, 0.0)) * delta) / delta ** 2
    gR = np.sin(np.sqrt(max(2 * (E - V_R) - kin, 0.0)) * delta) / delta ** 2
    return gL * gR * abs(G) ** 2


def test_transmission_uniform_closed_form():
    L = 20.0
    R2 = 9.5
    ENCUT = R2 * (2 * np.pi / L) ** 2 / 2
    nx, ny = O.planewave_basis(ENCUT, L, L)
    N_z = 40
    NG = 8
    Vrec = np.zeros((2 * NG + 1, 2 * NG + 1, N_z), dtype=np.complex128)
    a = O.Atomic(ETA=1e-6, LX=L, LY=L, LZ=N_z * 0.5, ENCUT=ENCUT)
    s = O.System(a, nx, ny, Vrec, NG, NG)
    kin = O.kinetic(nx, ny, L, L, 0.0, 0.0)
    for E in [0.05, 0.2, 0.45]:
        T = O.transmission_tau(s, E, 0)
        T2 = O.transmission_tau(s, E, 0, faithful_solver=False)
        Tc = sum(chain_T_closed_form(E, a.ETA, kin[i], a.LZ / N_z, N_z) for i in range(len(nx)))
        assert T == pytest.approx(Tc, rel=1e-9, abs=1e-12)
        assert T2 == pytest.approx(T, rel=1e-10, abs=1e-13)
        n_open = int(np.sum(2 * E - kin > 0))
        assert 0 <= T <= n_open + 1e-9


def test_transmission_mirror_symmetry_and_positivity():
    s, V = tiny_system(N_z=6, seed=11, eta=1e-7)
    # mirrored potential in z
    s2 = O.System(s.atomic, s.nx, s.ny, s.Vrec[:, :, ::-1].copy(), s.NGX, s.NGY)
    for E in [0.3, 0.8, 1.5]:
        t1 = O.transmission_tau(s, E, 0)
        t2 = O.transmission_tau(s2, E, 0)
        assert t1 >= 0
        assert t1 == pytest.approx(t2, rel=1e-8, abs=1e-12)


# ------------------------------------------------------------------ quadrature
@pytest.mark.parametrize("N", [3, 5, 6, 7, 8, 11, 12])
def test_simpson_exact_for_cubics(N):
    a, b = -0.3, 1.7
    xs = [O.value_simpson(i, N, a, b) for i in range(1, N + 1)]
    cs = [O.coefficient_simpson(i, N, a, b) for i in range(1, N + 1)]
    assert sum(cs) == pytest.approx(b - a, rel=1e-14)
    f = lambda x: 2 * x ** 3 - x ** 2 + 0.5 * x - 3
    F = lambda x: x ** 4 / 2 - x ** 3 / 3 + x ** 2 / 4 - 3 * x
    assert sum(c * f(x) for c, x in zip(cs, xs)) == pytest.approx(F(b) - F(a), rel=1e-12)
    assert xs[0] == a and xs[-1] == pytest.approx(b)


def test_simpson_quirk_N4_is_wrong():
    cs = [O.coefficient_simpson(i, 4, 0.0, 3.0) for i in range(1, 5)]
    assert sum(cs) != pytest.approx(3.0)   # reference quirk kept (i==1 branch wins)


def test_fermi():
    assert O.fermi_func(0.1, 0.1, 300.0) == 0.5
    x = 0.01
    assert O.fermi_func(0.1 + x, 0.1, 77.0) + O.fermi_func(0.1 - x, 0.1, 77.0) == pytest.approx(1.0)


# ------------------------------------------------------------------ stage 4: density / LDOS
def test_contour_circle_closure():
    """The semicircle part of Equilibrium_Density equals -(2/pi) times the real-axis
    integral of G over [circle_L, circle_R] (G analytic in the upper half plane)."""
    s, V = tiny_system(N_z=3, seed=5, eta=2e-3)
    a = s.atomic
    a.MU, a.TEMPERATURE = 0.6, 300.0
    a.GAP = 6 * O.K_B * a.TEMPERATURE
    min_V = -0.2
    N_circle = 301
    nodes = O.density_contour(a, N_circle, 11, min_V)
    circ = sum(w * O.diagonal_blocks(s, E, 0) for E, w in nodes[:N_circle])
    cl, cr = min_V - 1.0 / O.HARTREE, a.MU - a.GAP
    # dense real-axis Simpson of G(E + i eta)
    M = 4001
    tot = 0
    for i in range(1, M + 1):
        E = O.value_simpson(i, M, cl, cr)
        tot = tot + O.coefficient_simpson(i, M, cl, cr) * O.diagonal_blocks(s, E, 0)
    np.testing.assert_allclose(circ, -2 / np.pi * tot, rtol=2e-4, atol=2e-5 * np.abs(tot).max())


def test_gauss_legendre_nodes_match_numpy():
    """The host's Gauss-Legendre rule (own Newton iteration) against
    numpy.polynomial.legendre.leggauss, and exactness up to degree 2n - 1."""
    from negftrans_amd import negf
    for n in (1, 2, 3, 7, 16, 40, 64):
        x, w = negf.gauss_legendre(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.max(np.abs(x - xr)) < 1e-14 and np.max(np.abs(w - wr)) < 1e-13
        for deg in (0, 2 * n - 2, 2 * n - 1):
            exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
            assert np.sum(w * x ** deg) == pytest.approx(exact, abs=1e-13)


def test_gauss_legendre_contour_converges_to_simpson_reference():
    """Gauss-Legendre contour (extension): the host and the oracle give the same nodes,
    and the circle part converges to a dense (N = 301) Simpson reference faster than
    the reference's Simpson rule with a comparable node count."""
    from negftrans_amd import negf
    s, V = tiny_system(N_z=3, seed=5, eta=2e-3)
    a = s.atomic
    a.MU, a.TEMPERATURE = 0.6, 300.0
    a.GAP = 6 * O.K_B * a.TEMPERATURE
    min_V = -0.2
    ha = negf.Atomic(MU=a.MU, TEMPERATURE=a.TEMPERATURE, GAP=a.GAP, ETA=a.ETA, LX=a.LX, LY=a.LY, LZ=a.LZ,
                     ENCUT=a.ENCUT)
    z, w = negf.density_contour(ha, 12, 9, min_V, "gauss-legendre")
    nodes = O.density_contour(a, 12, 9, min_V, "gauss-legendre")
    np.testing.assert_allclose(z, [e for e, _ in nodes], rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(w, [q for _, q in nodes], rtol=1e-13, atol=1e-15)

    def circle(n, quad):
        nd = O.density_contour(a, n, 3, min_V, quad)[:n]
        return sum(q * O.diagonal_blocks(s, e, 0) for e, q in nd)
    ref = circle(301, "simpson")
    scale = np.abs(ref).max()
    e_gl = [np.abs(circle(n, "gauss-legendre") - ref).max() / scale for n in (8, 16, 24)]
    e_si = np.abs(circle(17, "simpson") - ref).max() / scale
    assert e_gl[0] > e_gl[1] > e_gl[2] and e_gl[2] < 1e-6 and e_gl[1] < e_si / 100


def test_neq_spectral_sum_rule():
    """Bias extension: the lead spectral blocks add up to the device spectral function,
    i (G - G^dagger) = G (Gamma_L + Gamma_R + 2 eta) G^dagger on every diagonal block,
    with Gamma from the trace's real-energy formula (exact up to O(eta) for eta -> 0).
    This pins the slice each lead sits on, the 2 Delta^2 scaling and the conjugation."""
    s, V = tiny_system(N_z=4, seed=12, eta=1e-10, V_L=0.01, V_R=-0.02)
    a = s.atomic
    E = 1.3
    assert np.sum(O.gamma_lead(E, a.V_L, s.nx, s.ny, a.LX, a.LY, a.LZ, s.N_z, 0, 0) > 0) >= 2
    AL = O.neq_spectral_blocks(s, E, 0, "L")
    AR = O.neq_spectral_blocks(s, E, 0, "R")
    D = O._blocks(s, 0, E + 1j * a.ETA)
    G = np.linalg.inv(O.assemble_dense(D)) * 2.0 * (a.LZ / s.N_z) ** 2
    N = s.N
    for k in range(s.N_z):
        Gk = G[k * N:(k + 1) * N, k * N:(k + 1) * N]
        lhs = 1j * (Gk - Gk.conj().T)
        rows = G[k * N:(k + 1) * N]
        rhs = AL[k] + AR[k] + 2 * a.ETA * rows @ rows.conj().T
        np.testing.assert_allclose(rhs, lhs, rtol=0, atol=1e-6 * np.abs(lhs).max())
        # each part is Hermitian positive semidefinite
        np.testing.assert_allclose(AL[k], AL[k].conj().T, atol=1e-12 * np.abs(AL[k]).max())
        assert np.linalg.eigvalsh(AL[k]).min() > -1e-10 * np.abs(AL[k]).max()


def test_neq_bias_window_and_zero_bias_limit():
    from negftrans_amd import negf
    s, V = tiny_system(N_z=3, seed=5, eta=2e-3)
    a = s.atomic
    a.MU, a.TEMPERATURE = 0.6, 300.0
    a.GAP = 6 * O.K_B * a.TEMPERATURE
    VDS = 0.01
    for quad in ("simpson", "gauss-legendre"):
        E, w = O.bias_window(a, VDS, 61, quad)
        ha = negf.Atomic(MU=a.MU, TEMPERATURE=a.TEMPERATURE, GAP=a.GAP)
        Eh, wh = negf.bias_window(ha, VDS, 61, quad)
        np.testing.assert_allclose(Eh, E, rtol=1e-15, atol=1e-15)
        np.testing.assert_allclose(wh, w, rtol=1e-13, atol=1e-17)
        # integral of f_L - f_R over the window is VDS up to the exp(-6) Fermi tails
        assert np.pi * np.sum(w.imag) == pytest.approx(VDS, rel=0, abs=1e-2 * O.K_B * a.TEMPERATURE)
    DM0 = O.density_matrix_bias(s, 9, 7, -0.2, 0.0, 5)
    np.testing.assert_array_equal(DM0, O.density_matrix(s, 9, 7, -0.2))


def test_neq_closed_right_lead_equals_equilibrium_at_mu_L():
    """Bias extension, end to end: with the right lead closed at every window energy
    (V_R = 50 Hartree, so Gamma_R = 0), G Gamma_L G^dagger is the whole spectral function up
    to O(eta), and the equilibrium contour at mu_R plus the bias window must give the
    equilibrium density at mu_L.  Compared on the anti-Hermitian part (DM - DM^H) / 2i,
    which the real-space density reads.  GAP = 36 kT removes the Fermi-tail truncation.
    A wrong sign of f_L - f_R, a dropped 1/pi, a wrong Gamma scale or lead slice, or a
    missing (2 Delta^2)^2 each fail it (the window term is 1.8 % of the density here,
    the bar 1e-5; measured 1e-6)."""
    s, V = tiny_system(N_z=3, seed=5, eta=1e-5, V_R=50.0)
    a = s.atomic
    a.MU, a.TEMPERATURE = 0.9, 300.0
    a.GAP = 36 * O.K_B * a.TEMPERATURE
    VDS = 0.01

    def ah(M):
        return (M - np.conj(np.transpose(M, (0, 2, 1)))) / 2j

    ref = ah(O.density_matrix(s, 40, 61, float(V.min()), "gauss-legendre"))
    got = ah(O.density_matrix_bias(s, 40, 61, float(V.min()), VDS, 61, "gauss-legendre"))
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
    import dataclasses
    s_R = dataclasses.replace(s, atomic=dataclasses.replace(a, MU=a.MU - VDS))
    eq_R = ah(O.density_matrix(s_R, 40, 61, float(V.min()), "gauss-legendre"))
    assert np.linalg.norm(eq_R - ref) > 1e-2 * np.linalg.norm(ref)


def test_ldos_matrix_vs_dense():
    s, V = tiny_system(N_z=4, seed=8, eta=1e-3)
    E = 0.9
    LM = O.ldos_matrix(s, E)
    a = s.atomic
    D = O._blocks(s, 0, E + 1j * a.ETA)
    full = np.linalg.inv(O.assemble_dense(D)) * 2 * (a.LZ / s.N_z) ** 2
    N = s.N
    for k in range(s.N_z):
        np.testing.assert_allclose(LM[k], -full[k * N:(k + 1) * N, k * N:(k + 1) * N] / np.pi,
                                   rtol=1e-10, atol=1e-13)


def test_pw_to_real_literal_vs_fast():
    L = 6.0
    ENCUT = 4.5 * (2 * np.pi / L) ** 2 / 2
    nx, ny = O.planewave_basis(ENCUT, L, L)
    N = len(nx)
    N_x, N_y = 6, 8
    M = rng(9).standard_normal((N, N)) + 1j * rng(10).standard_normal((N, N))
    np.testing.assert_allclose(O.pw_to_real(M, nx, ny, N_x, N_y),
                               O.pw_to_real_literal(M, nx, ny, N_x, N_y), rtol=1e-12, atol=1e-12)
    out = O.pw_to_real(np.eye(N), nx, ny, N_x, N_y)
    np.testing.assert_allclose(out, np.full((N_x, N_y), N), atol=1e-12)


def test_ldos_grid_converter():
    assert O.ldos_grid_converter(3, 0, 0.5, 2.0) == (3, 2.0)
    assert O.ldos_grid_converter(3, -1, 0.5, 2.0) == (1, 1.0)
    assert O.ldos_grid_converter(3, 3, 0.5, 2.0) == (1, 2.0)
    assert O.ldos_grid_converter(3, 4, 0.5, 2.0) == (1, 0.0)
    assert O.ldos_size(0, 7) == 7 and O.ldos_size(-1, 7) == 1 and O.ldos_size(2, 7) == 1


def test_lead_decimation_matches_lattice_root():
    """Decimated end-slice term (extension) against the analytic surface Green's function of
    the chain tridiag(1, d, 1): g_s solves g^2 - d g + 1 = 0 with |g_s| < 1 (the decaying,
    retarded root), term = -g_s.  Open, closed and above-band channels, eta > 0."""
    s, V = tiny_system(N_z=6, seed=3, eta=1e-7)
    a = s.atomic
    delta = a.LZ / s.N_z
    seen = np.zeros(3, dtype=bool)          # below band, open, above band
    for E, Vl in ((0.3, 0.0), (1.7, 0.2), (2.5, -0.1)):
        term, it = O.lead_decimation(E + 1j * a.ETA, Vl, s.nx, s.ny, a.LX, a.LY, delta, 0.0, 0.0)
        kin = O.kinetic(s.nx, s.ny, a.LX, a.LY, 0.0, 0.0)
        d = 2 * delta ** 2 * (E + 1j * a.ETA - Vl - kin / 2) - 2
        r = np.sqrt(d * d - 4 + 0j)
        g1, g2 = (d - r) / 2, (d + r) / 2
        gs = np.where(np.abs(g1) < np.abs(g2), g1, g2)
        assert (it > 0).all() and it.max() < 200
        np.testing.assert_allclose(term, -gs, rtol=0, atol=1e-11)
        # open channels: |term| = 1 up to eta, and the lattice dispersion holds
        op = O.lattice_open(E, Vl, s.nx, s.ny, a.LX, a.LY, delta, 0.0, 0.0)
        x = delta ** 2 * (E - Vl - kin / 2)
        seen |= [(x <= 0).any(), op.any(), (x >= 2).any()]
        np.testing.assert_allclose(np.abs(term[op]), 1.0, atol=1e-5)
        np.testing.assert_allclose(term[op].real, 1 - delta ** 2 * (E - Vl - kin[op] / 2), atol=1e-5)
        assert (np.abs(term[~op]) < 1).all()
    assert seen.all()


def test_lead_decimation_uniform_wire_transmits_one_per_open_channel():
    """V = 0 everywhere: with the lattice lead, every open channel transmits 1 (the device
    is the lead continued).  The reference's continuum closed form gives less (mismatch
    between the continuum q and the lattice dispersion), which the second check shows."""
    import dataclasses
    s, V = tiny_system(N_z=6, seed=3, eta=1e-7)
    s.Vrec[:] = 0
    a = s.atomic
    sd = dataclasses.replace(s, lead="decimation")
    delta = a.LZ / s.N_z
    for E in (0.3, 0.9, 1.7, 2.5):
        n_open = int(np.sum(O.lattice_open(E, 0.0, s.nx, s.ny, a.LX, a.LY, delta, 0.0, 0.0)))
        T = O.transmission(sd, [E])[0]
        assert abs(T - n_open) <= 1e-5 * max(n_open, 1)
        assert abs(O.transmission(s, [E])[0] - n_open) > 4e-4


def test_nonlocal_potential_literal_sum():
    """NonLocalPotential_RealToPlanewave restated through FFTs against the literal four-fold
    sum (1/(N_x N_y)^2) sum V(x,y,x',y') exp(-2 pi i (p x/N_x + q y/N_y - p' x'/N_x - q' y'/N_y))."""
    Nx, Ny, Rx, Ry = 6, 4, 2, 1
    r = rng(31)
    V = r.standard_normal((Nx, Ny, Nx, Ny)) + 1j * r.standard_normal((Nx, Ny, Nx, Ny))
    got = O.nonlocal_potential_r2p(V, Rx, Ry)
    x = np.arange(Nx)[:, None, None, None]
    y = np.arange(Ny)[None, :, None, None]
    xp = np.arange(Nx)[None, None, :, None]
    yp = np.arange(Ny)[None, None, None, :]
    for p in range(-Rx, Rx + 1):
        for q in range(-Ry, Ry + 1):
            for pp in range(-Rx, Rx + 1):
                for qq in range(-Ry, Ry + 1):
                    ph = np.exp(-2j * np.pi * (p * x / Nx + q * y / Ny - pp * xp / Nx - qq * yp / Ny))
                    want = np.sum(V * ph) / (Nx * Ny) ** 2
                    assert abs(got[p + Rx, q + Ry, pp + Rx, qq + Ry] - want) < 1e-13


def test_nonlocal_of_a_local_potential():
    """V(r, r') = delta(r - r') v(r): the non-local transform is the local one of the
    difference frequency, V(p,q,p',q') = v_hat(p - p', q - q') / (N_x N_y), with v_hat from
    local_potential_r2p (pinned by its own closed forms)."""
    Nx, Ny, R = 10, 10, 2      # the local window +-4 needs NGY + 1 + N_y / 2 <= N_y
    v = rng(32).standard_normal((Nx, Ny))
    V = np.zeros((Nx, Ny, Nx, Ny), dtype=np.complex128)
    for i in range(Nx):
        for j in range(Ny):
            V[i, j, i, j] = v[i, j]
    got = O.nonlocal_potential_r2p(V, R, R)
    loc = O.local_potential_r2p(v, 2 * R, 2 * R)
    for p in range(-R, R + 1):
        for q in range(-R, R + 1):
            for pp in range(-R, R + 1):
                for qq in range(-R, R + 1):
                    want = loc[p - pp + 2 * R, q - qq + 2 * R] / (Nx * Ny)
                    assert abs(got[p + R, q + R, pp + R, qq + R] - want) < 1e-14


# ------------------------------------------------------------------ output units
def test_density_real_charge_sum_rule():
    """density_real (src/main.f90:341-349) against the number of electrons: the density
    summed over the N_x x N_y x N_z grid times the cell volume (LX/N_x)(LY/N_y)(LZ/N_z)
    equals Im Tr of the slice's density matrix, slice by slice.  A wrong N_z factor, a
    normalised transform or a missing 1/(LX LY LZ) fails it; so does a transform that
    does not reduce to the trace (every basis difference n_i - n_j != 0 integrates to 0)."""
    L = 7.0
    ENCUT = 9.5 * (2 * np.pi / L) ** 2 / 2
    nx, ny = O.planewave_basis(ENCUT, L, L)
    N, N_z, N_x, N_y = len(nx), 4, 14, 10
    LXa, LYa, LZa = 23.0, 17.0, 31.0                 # Angstrom, unequal on purpose
    r = rng(12)
    DM = r.standard_normal((N_z, N, N)) + 1j * r.standard_normal((N_z, N, N))
    dens = O.density_real(DM, nx, ny, N_x, N_y, LXa, LYa, LZa)
    dV = (LXa / N_x) * (LYa / N_y) * (LZa / N_z)
    for k in range(N_z):
        assert dens[:, :, k].sum() * dV == pytest.approx(np.trace(DM[k]).imag, rel=1e-12, abs=1e-12)
    # one diagonal element: a uniform density Im(M_ii) / slice volume
    M = np.zeros((1, N, N), dtype=np.complex128)
    M[0, 3, 3] = 0.25j
    d1 = O.density_real(M, nx, ny, N_x, N_y, LXa, LYa, LZa)
    np.testing.assert_allclose(d1, 0.25 / (LXa * LYa * LZa), rtol=1e-12)


def test_ldos_units_against_lattice_dos():
    """ldos_matrix + ldos_reduce (src/negf.f90:176-236, src/main.f90:510-525,
    src/tools.f90:18-40) against the textbook density of states of the device's own
    lattice.  With V = 0 and the decimated (lattice) leads the wire is an infinite chain
    per plane wave, tridiag(1, d_i, 1) with d_i = 2 Delta^2 (E - |G_i|^2 / 2) - 2, whose
    site Green's function is -i / sqrt(4 - d_i^2) in the band; G = 2 Delta^2 A^-1 makes
    the DOS per slice sum_i 2 Delta^2 / (pi sqrt(4 - d_i^2)) per Hartree.  Reduced with
    LDOS_GRID = (-1, -1, 0) (integrate x and y, keep z) the reference's output is that
    per unit length (/ Delta in Angstrom) and per eV (/ hartree); summed over all three
    axes (-1, -1, -1) it is the DOS of the whole wire per eV."""
    import dataclasses
    s, _ = tiny_system(N_z=7, seed=5, eta=1e-8)
    s.Vrec[:] = 0
    sd = dataclasses.replace(s, lead="decimation")
    a = s.atomic
    delta = a.LZ / s.N_z                               # bohr
    half_g2 = 0.5 * O.kinetic(s.nx, s.ny, a.LX, a.LY, 0.0, 0.0)     # |G_i|^2 / 2
    N_x = N_y = 8
    LXa, LYa, LZa = a.LX * O.A_0, a.LY * O.A_0, a.LZ * O.A_0
    for E in (0.35, 1.3, 2.2):
        d = 2 * delta ** 2 * (E - half_g2) - 2
        inband = np.abs(d) < 2
        assert inband.any()
        dos_slice = np.sum(2 * delta ** 2 / (np.pi * np.sqrt(4 - d[inband] ** 2)))   # per Hartree
        LM = O.ldos_matrix(sd, E)
        z = O.ldos_reduce(LM, s.nx, s.ny, N_x, N_y, LXa, LYa, LZa, (-1, -1, 0))[0, 0, :]
        np.testing.assert_allclose(z, dos_slice / (delta * O.A_0) / O.HARTREE, rtol=1e-5)
        tot = O.ldos_reduce(LM, s.nx, s.ny, N_x, N_y, LXa, LYa, LZa, (-1, -1, -1))[0, 0, 0]
        assert tot == pytest.approx(s.N_z * dos_slice / O.HARTREE, rel=1e-5)


@pytest.mark.parametrize("L,R2", [(400.0, 19.0), (400.0, 100.5), (400.0, 256.5), (123.0, 7.3), (55.5, 40.2)])
def test_ngx_window_holds_every_basis_difference(L, R2):
    """The NGX/NGY auto rule (src/main.f90:94-100) against what the window is for:
    V_reciprocal_all(-NGX:NGX, -NGY:NGY) must hold every plane-wave difference
    n_i - n_j that Hamiltonian_construction reads (src/grid.f90:252), i.e.
    NGX >= 2 max |n_x| over the basis, found here by enumerating the basis."""
    L_ang = L
    Lb = L_ang / O.A_0
    ENCUT_eV = R2 * (2 * np.pi / Lb) ** 2 / 2 * O.HARTREE
    nx, ny = O.planewave_basis(ENCUT_eV / O.HARTREE, Lb, Lb)
    NGX, NGY = O.ngx_auto(0, 0, L_ang, L_ang, ENCUT_eV)
    assert NGX >= 2 * np.max(np.abs(nx)) and NGY >= 2 * np.max(np.abs(ny))
    assert NGX <= 2 * np.max(np.abs(nx)) + 3          # and no larger than the rounding needs
