"""CPU FP64 oracle: a restatement of NEGFtrans's RGF energy loop in numpy/LAPACK.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (negftrans_amd/) imports this
module; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
use it, and only as the checker / the timed CPU baseline.

Every function cites the reference routine it restates (paths relative to the
NEGFtrans source tree).  Atomic units throughout (Hartree, bohr, m_e = hbar = 1),
exactly as the reference after its unit conversion (src/main.f90:182-205).

Third-party arithmetic: the reference binds LAPACK zgetrf/zgetri/ilaenv and BLAS
zgemm through `-llapack -lblas` with no pinned version (build/makefile:10) and
FFTW3 for the 2-D transforms (src/grid.f90:5).  Here `inverse` calls the LAPACK
zgetrf/zgetri and `matrix_product` the BLAS zgemm that ship with scipy
(OpenBLAS 0.3.28), i.e. the same routines; the transforms use numpy's FFT,
which computes the same unnormalised DFT (FFTW_FORWARD = e^{-2 pi i ...}).

Pinning status of each function (what checks it in tests/test_oracle_*.py):
  planewave_basis / findsize   Gauss-circle counts, ordering rule          pinned
  kpoint_mesh                  Monkhorst-Pack closed form, weight sum       pinned
  local_potential_r2p          closed forms (constant, cosines), hermiticity,
                               literal DFT sum on tiny grids               pinned
  hamiltonian / e_minus_h      brute-force assembly of the 3-D operator     pinned
  inverse / matrix_product     numpy.linalg.inv / A@B residuals             pinned
  tri_solver (D, F, L)         dense (N*N_z)^2 inversion on tiny systems    pinned
  transmission_solver          V=0 decoupled chains (1-D closed form),
                               T>=0, mirror symmetry                        pinned
  simpson / fermi              exact integrals of cubics, (b-a) sum         pinned
  equilibrium_density          contour closure vs dense real-axis integral  pinned (closure)
  local_density_of_state       dense inverse of the assembled operator      pinned
  pw_to_real                   literal (NxNy)^2 tensor form on tiny grids   pinned
  V-hat window truncation (src/main.f90:242-246) and the Gamma_L/Gamma_R
  placement (src/grid.f90:362-383) follow the reference text line by line;
  no reference run exists to pin their composition ("parity unpinned" for
  those two conventions beyond the line-by-line restatement).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg.blas as _blas
import scipy.linalg.lapack as _lapack

# ---------------------------------------------------------------- constants
# src/global.f90:24-28
HARTREE = 27.211386245988            # eV
K_B = 8.617333262e-5 / HARTREE       # Hartree / K
A_0 = 0.529177210903                 # Angstrom
PI = 3.141592653589793


@dataclass
class Atomic:
    """t_parameters in atomic units (src/global.f90:5-7, src/main.f90:182-198)."""
    V_L: float = 0.0
    V_R: float = 0.0
    MU: float = 5.568 / HARTREE
    ETA: float = 1e-5
    TEMPERATURE: float = 20.0
    LX: float = 0.0
    LY: float = 0.0
    LZ: float = 0.0
    ENCUT: float = 100.0 / HARTREE
    GAP: float = 6.0 * K_B * 20.0
    VDS: float = 0.0
    TRANSMISSION_GRID: tuple = (0.0, 0.0)
    LDOS_ENERGY_GRID: tuple = (0.0, 0.0)


def atomic_from_input(V_L=0.0, V_R=0.0, MU=5.568, ETA=1e-5, TEMPERATURE=20.0,
                      LX=0.0, LY=0.0, LZ=0.0, ENCUT=100.0, GAP=6.0, VDS=0.0,
                      TRANSMISSION_GRID=(0.0, 0.0, 0.0),
                      LDOS_ENERGY_GRID=(0.0, 0.0, 0.0)) -> Atomic:
    """Unit conversion of the INPUT namelist (src/main.f90:182-198). eV/Angstrom in,
    Hartree/bohr out; ETA is already Hartree; GAP is in k_B*T units."""
    return Atomic(
        V_L=V_L / HARTREE, V_R=V_R / HARTREE, MU=MU / HARTREE, ETA=ETA,
        TEMPERATURE=TEMPERATURE, LX=LX / A_0, LY=LY / A_0, LZ=LZ / A_0,
        ENCUT=ENCUT / HARTREE, GAP=GAP * K_B * TEMPERATURE, VDS=VDS / HARTREE,
        TRANSMISSION_GRID=(TRANSMISSION_GRID[0] / HARTREE, TRANSMISSION_GRID[1] / HARTREE),
        LDOS_ENERGY_GRID=(LDOS_ENERGY_GRID[0] / HARTREE, LDOS_ENERGY_GRID[1] / HARTREE))


def ngx_auto(NGX: int, NGY: int, LX_ang: float, LY_ang: float, ENCUT_eV: float):
    """NGX/NGY auto rule (src/main.f90:94-100).  Quirk kept: the NGX test uses LY."""
    sx = math.ceil((LX_ang / A_0) / PI * math.sqrt(ENCUT_eV / HARTREE / 2) * 2)
    sy = math.ceil((LY_ang / A_0) / PI * math.sqrt(ENCUT_eV / HARTREE / 2) * 2)
    if NGX <= sy:
        NGX = sx + 1
    if NGY <= sy:
        NGY = sy + 1
    return NGX, NGY


# ---------------------------------------------------------------- basis, k-mesh
def _nmax(L, ENCUT):
    return math.ceil(L / PI * math.sqrt(ENCUT / 2))


def planewave_basis_findsize(ENCUT: float, Lx: float, Ly: float) -> int:
    """src/grid.f90:387-408."""
    return len(planewave_basis(ENCUT, Lx, Ly)[0])


def planewave_basis(ENCUT: float, Lx: float, Ly: float):
    """src/grid.f90:410-431: all (i, j) with ((2 pi i/Lx)^2 + (2 pi j/Ly)^2)/2 <= ENCUT,
    j outer ascending, i inner ascending.  Returns int32 arrays nx, ny."""
    nxm, nym = _nmax(Lx, ENCUT), _nmax(Ly, ENCUT)
    nx, ny = [], []
    for j in range(-nym, nym + 1):
        for i in range(-nxm, nxm + 1):
            if ((2 * i * PI / Lx) ** 2 + (2 * j * PI / Ly) ** 2) / 2 <= ENCUT:
                nx.append(i)
                ny.append(j)
    return np.array(nx, dtype=np.int32), np.array(ny, dtype=np.int32)


def kpoint_mesh(NKX: int, NKY: int, Lx: float, Ly: float):
    """src/grid.f90:433-448 (Monkhorst-Pack). Returns arrays kx, ky, weight."""
    kx, ky, w = [], [], []
    for j in range(1, NKY + 1):
        for i in range(1, NKX + 1):
            kx.append((i - (1 + NKX) / 2.0) / NKX * (2 * PI / Lx))
            ky.append((j - (1 + NKY) / 2.0) / NKY * (2 * PI / Ly))
            w.append(1.0 / (NKX * NKY))
    return np.array(kx), np.array(ky), np.array(w)


def kinetic(nx, ny, Lx, Ly, kx, ky):
    """|G_i + k|^2 as used in src/grid.f90:258,285-288,356-359 (without the 1/2)."""
    return (2 * nx * PI / Lx + kx) ** 2 + (2 * ny * PI / Ly + ky) ** 2


# ---------------------------------------------------------------- stage 1: V-hat
def local_potential_r2p(V_slice: np.ndarray, NGX: int, NGY: int) -> np.ndarray:
    """LocalPotential_RealToPlanewave (src/grid.f90:12-69) for one z-slice.

    V_slice is (N_x, N_y) real (Fortran index order x, y).  Returns the array
    V_reciprocal(-NGX:NGX, -NGY:NGY) stored as [p + NGX, q + NGY]:
        V(p, q) = (1/(N_x N_y)) sum_{x,y} V(x, y) exp(-2 pi i (p x/N_x + q y/N_y))
    for p = 0..NGX, all q in the window (src/grid.f90:58-62, the (-1)^(j-1) factor
    at :46-50 shifts q by N_y/2), and p < 0 by conjugate symmetry (:64-68)."""
    N_x, N_y = V_slice.shape
    if 2 * NGX > N_x or 2 * NGY > N_y:
        raise ValueError("size of V_reciprocal minus one larger than size of V_real")
    if NGX > N_x // 2 or NGY + 1 + N_y // 2 > N_y:
        raise ValueError("V_reciprocal window does not fit the real-space grid")
    F = np.fft.fft2(V_slice) / (N_x * N_y)
    out = np.zeros((2 * NGX + 1, 2 * NGY + 1), dtype=np.complex128)
    for p in range(0, NGX + 1):
        for q in range(-NGY, NGY + 1):
            out[p + NGX, q + NGY] = F[p % N_x, q % N_y]
    for p in range(-NGX, 0):
        for q in range(-NGY, NGY + 1):
            out[p + NGX, q + NGY] = np.conj(out[-p + NGX, -q + NGY])
    return out


def nonlocal_potential_r2p(V_real: np.ndarray, Rx: int, Ry: int) -> np.ndarray:
    """NonLocalPotential_RealToPlanewave (src/grid.f90:137-234), never called by the
    reference program.  V_real(N_x, N_y, N_x, N_y) complex, Fortran index order
    (x, y, x', y').  The steps of the reference: multiply by (-1)^(i+j-ii-jj)
    (:171-179), unnormalised backward 2-D DFT over (x', y') (:184-201), forward 2-D DFT
    over (x, y) (:203-219), divide by (N_x N_y)^2 and keep frequencies -R..R of every
    index, frequency f at 0-based position f + N/2 (:222-232).  Returns an array
    indexed [p + Rx, q + Ry, p' + Rx, q' + Ry].  At |f| = N/2 the reference reads past
    its work array (index N + 1); this restatement takes the periodic value."""
    N_x, N_y = V_real.shape[:2]
    if 2 * Rx > N_x or 2 * Ry > N_y:
        raise ValueError("size of V_reciprocal minus one larger than size of V_real")
    x = np.arange(N_x)[:, None, None, None]
    y = np.arange(N_y)[None, :, None, None]
    xp = np.arange(N_x)[None, None, :, None]
    yp = np.arange(N_y)[None, None, None, :]
    W = V_real * (-1.0) ** (x + y - xp - yp)
    W = np.fft.ifft2(W, axes=(2, 3)) * (N_x * N_y)
    W = np.fft.fft2(W, axes=(0, 1))
    ix = (np.arange(-Rx, Rx + 1) + N_x // 2) % N_x
    iy = (np.arange(-Ry, Ry + 1) + N_y // 2) % N_y
    return W[np.ix_(ix, iy, ix, iy)] / float(N_x * N_y) ** 2


def v_reciprocal_all(V_real: np.ndarray, NGX: int, NGY: int, full_window: bool = False):
    """The V_reciprocal_all(-NGX:NGX, -NGY:NGY, N_z) table built by the reference's
    main program (src/main.f90:217-249): per slice the transform above, but only
    |p| <= NGX/2, |q| <= NGY/2 are copied (:242-246); the rest of the table is
    never written by the reference (a fresh allocation, zero here).
    full_window=True copies the whole +-NGX, +-NGY window (extension)."""
    N_x, N_y, N_z = V_real.shape
    out = np.zeros((2 * NGX + 1, 2 * NGY + 1, N_z), dtype=np.complex128)
    hx, hy = (NGX, NGY) if full_window else (NGX // 2, NGY // 2)
    for k in range(N_z):
        vr = local_potential_r2p(V_real[:, :, k], NGX, NGY)
        out[NGX - hx:NGX + hx + 1, NGY - hy:NGY + hy + 1, k] = \
            vr[NGX - hx:NGX + hx + 1, NGY - hy:NGY + hy + 1]
    return out


# ---------------------------------------------------------------- blocks
def hamiltonian(nx, ny, Lx, Ly, kx, ky, Vrec_all, NGX, NGY):
    """Hamiltonian_construction (src/grid.f90:236-262):
    H_k(i, j) = V(nx_i - nx_j, ny_i - ny_j, k) + delta_ij |G_i + k|^2 / 2."""
    dx = (nx[:, None] - nx[None, :]) + NGX
    dy = (ny[:, None] - ny[None, :]) + NGY
    H = np.ascontiguousarray(np.moveaxis(Vrec_all[dx, dy, :], 2, 0))  # (N_z, N, N)
    kin = kinetic(nx, ny, Lx, Ly, kx, ky) / 2
    idx = np.arange(len(nx))
    H[:, idx, idx] += kin
    return H


def lead_q(energy: complex, V_lead: float, nx, ny, Lx, Ly, kx, ky):
    """QR/QT of src/grid.f90:284-289: principal complex sqrt."""
    return np.sqrt(2 * (energy - V_lead) - kinetic(nx, ny, Lx, Ly, kx, ky) + 0j)


def lead_decimation(energy: complex, V_lead: float, nx, ny, Lx, Ly, delta, kx, ky,
                    tol: float = 1e-14, max_iter: int = 200):
    """Extension (no reference counterpart; the reference adds exp(i q Delta) with the
    continuum q, src/grid.f90:284-289,323-331): the end-slice term of a semi-infinite flat
    lead on the same finite-difference lattice as the device, by Sancho-Rubio decimation.

    The scaled lead is the chain tridiag(1, d_i, 1) per basis function i, with
    d_i = 2 Delta^2 (E - V_lead - |G_i + k|^2 / 2) - 2.  Doubling step, all scalars:
        g = 1 / e;  e_s -= a g b;  e -= a g b + b g a;  a <- -a g a;  b <- -b g b
    from e_s = e = d, a = b = 1, until |re a| + |im a| < tol.  The surface Green's
    function is 1 / e_s and the device's end block gets -1 / e_s (for an open channel
    exp(i k Delta) with cos(k Delta) = 1 - Delta^2 (E - V_lead - |G + k|^2 / 2)).
    Returns (term (N,), iterations (N,); -1 where max_iter was reached)."""
    d = 2 * delta ** 2 * (energy - V_lead - kinetic(nx, ny, Lx, Ly, kx, ky) / 2) - 2 + 0j
    e_s, e = d.copy(), d.copy()
    a, b = np.ones_like(d), np.ones_like(d)
    it = np.full(d.shape, -1)
    done = np.zeros(d.shape, dtype=bool)
    for n in range(1, max_iter + 1):
        g = 1.0 / e
        agb, bga = a * g * b, b * g * a
        act = ~done
        e_s = np.where(act, e_s - agb, e_s)
        e = np.where(act, e - agb - bga, e)
        a, b = np.where(act, -a * g * a, a), np.where(act, -b * g * b, b)
        new = act & (np.abs(a.real) + np.abs(a.imag) < tol)
        it[new] = n
        done |= new
        if done.all():
            break
    return -1.0 / e_s, it


def lattice_open(energy: float, V_lead: float, nx, ny, Lx, Ly, delta, kx, ky):
    """Open channels of the lattice lead at real energy: 0 < Delta^2 (E - V - |G+k|^2/2) < 2."""
    x = delta ** 2 * (energy - V_lead - kinetic(nx, ny, Lx, Ly, kx, ky) / 2)
    return (x > 0) & (x < 2)


def lead_terms(energy: complex, nx, ny, Lx, Ly, Lz, N_z, kx, ky, V_L, V_R, lead: str = "closed-form"):
    """The two end-slice terms added to D_1 and D_{N_z}."""
    delta = Lz / N_z
    if lead == "decimation":
        sL, itL = lead_decimation(energy, V_L, nx, ny, Lx, Ly, delta, kx, ky)
        sR, itR = lead_decimation(energy, V_R, nx, ny, Lx, Ly, delta, kx, ky)
        if (itL < 0).any() or (itR < 0).any():
            raise ValueError("lead decimation did not converge (eta = 0?)")
        return sL, sR
    if lead != "closed-form":
        raise ValueError(f"unknown lead {lead!r}")
    return (np.exp(1j * lead_q(energy, V_L, nx, ny, Lx, Ly, kx, ky) * delta),
            np.exp(1j * lead_q(energy, V_R, nx, ny, Lx, Ly, kx, ky) * delta))


def e_minus_h(H, energy: complex, nx, ny, Lx, Ly, Lz, kx, ky, V_L, V_R, lead: str = "closed-form"):
    """E_minus_H_construction (src/grid.f90:264-333).  H is (N_z, N, N).
    D_k = 2 Delta^2 (E I - H_k) - 2 I, plus exp(i q Delta) on the first and last
    slice diagonals (closed-form lead self-energy; lead="decimation" adds the lattice
    lead's term from lead_decimation instead)."""
    N_z, N, _ = H.shape
    delta = Lz / N_z
    sL, sR = lead_terms(energy, nx, ny, Lx, Ly, Lz, N_z, kx, ky, V_L, V_R, lead)
    idx = np.arange(N)
    D = -H
    D[:, idx, idx] += energy
    D *= (2 * delta ** 2)
    D[:, idx, idx] -= 2.0
    D[0, idx, idx] += sL
    D[N_z - 1, idx, idx] += sR
    return D


# ---------------------------------------------------------------- stage 3: RGF
def inverse(M: np.ndarray) -> np.ndarray:
    """inverse (src/math_kernel.f90:8-34): copy, zgetrf, zgetri; INFO ignored."""
    lu, piv, _ = _lapack.zgetrf(np.array(M, dtype=np.complex128, order="F"))
    inv, _ = _lapack.zgetri(lu, piv)
    return inv


def matrix_product(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """matrix_product (src/math_kernel.f90:36-42): zgemm('N','N'), alpha=1, beta=0."""
    return _blas.zgemm(1.0, A, B)


def tri_solver(G_inv_D: np.ndarray, DFL: str):
    """GreensFunction_tri_solver (src/math_kernel.f90:44-165).

    G_inv_D is (N_z, N, N) (slice index first here).  Off-diagonal blocks are I.
    Returns dict with 'diagonal', 'first_column', 'last_column' (N_z, N, N),
    zero where not requested (src/math_kernel.f90:75-77)."""
    N_z, N, _ = G_inv_D.shape
    if_D = any(c in DFL for c in "Dd")
    if_F = any(c in DFL for c in "Ff")
    if_L = any(c in DFL for c in "Ll")
    out = {k: np.zeros((N_z, N, N), dtype=np.complex128)
           for k in ("diagonal", "first_column", "last_column")}
    I = np.eye(N, dtype=np.complex128)
    gr = [None] * (N_z - 1)  # generator_r(:, :, 1..N_z-1) -> gr[0..N_z-2]
    if if_D or if_F:
        gr[N_z - 2] = inverse(G_inv_D[N_z - 1])                       # :82
        for k in range(2, N_z):                                        # :83-90
            gr[N_z - k - 1] = inverse(G_inv_D[N_z - k] - gr[N_z - k])
    if if_D:                                                           # :94-109
        out["diagonal"][0] = inverse(G_inv_D[0] - gr[0])
        for k in range(2, N_z + 1):
            Mt = matrix_product(out["diagonal"][k - 2], gr[k - 2]) + I
            out["diagonal"][k - 1] = matrix_product(gr[k - 2], Mt)
    if if_F:                                                           # :112-129
        out["first_column"][0] = inverse(G_inv_D[0] - gr[0])
        for k in range(2, N_z + 1):
            out["first_column"][k - 1] = -matrix_product(gr[k - 2], out["first_column"][k - 2])
    if if_L:                                                           # :132-162
        gl = [None] * (N_z - 1)
        gl[0] = inverse(G_inv_D[0])
        for k in range(2, N_z):
            gl[k - 1] = inverse(G_inv_D[k - 1] - gl[k - 2])
        out["last_column"][N_z - 1] = inverse(G_inv_D[N_z - 1] - gl[N_z - 2])
        for k in range(1, N_z):
            out["last_column"][N_z - k - 1] = -matrix_product(gl[N_z - k - 1],
                                                              out["last_column"][N_z - k])
    return out


def rgf_first_column_last_slice(G_inv_D: np.ndarray) -> np.ndarray:
    """Only first_column(:, :, N_z) of tri_solver(..., 'F') with O(N^2) state: the same
    backward generators g_m as src/math_kernel.f90:82-90, with the product
    g_{N_z} g_{N_z-1} ... g_2 accumulated during the backward pass instead of the
    forward recursion at :120-128 (same matrices, different association)."""
    N_z = G_inv_D.shape[0]
    g = inverse(G_inv_D[N_z - 1])
    Q = g.copy()
    for m in range(N_z - 1, 1, -1):
        g = inverse(G_inv_D[m - 1] - g)
        Q = matrix_product(Q, g)
    G11 = inverse(G_inv_D[0] - g)
    if N_z == 1:
        raise ValueError("N_z must be >= 2")
    sign = -1.0 if (N_z - 1) % 2 else 1.0
    return sign * matrix_product(Q, G11)


# ---------------------------------------------------------------- stage 4: trace
def lead_gammas(energy: float, nx, ny, Lx, Ly, Lz, N_z, kx, ky, V_L, V_R, lead: str = "closed-form",
                eta: float = 0.0):
    """Trace weights of the two leads.  Closed form (src/grid.f90:355-360): sin(Q Delta) /
    Delta^2 at real E, closed channels 0.  Decimation (extension): Im(term) / Delta^2 of
    the decimated end-slice term at E + i eta on the lattice-open channels, 0 elsewhere
    (Gamma = i (Sigma - Sigma^dagger) with Sigma = -term / (2 Delta^2))."""
    delta = Lz / N_z
    if lead == "decimation":
        sL, sR = lead_terms(energy + 1j * eta, nx, ny, Lx, Ly, Lz, N_z, kx, ky, V_L, V_R, lead)
        oL = lattice_open(energy, V_L, nx, ny, Lx, Ly, delta, kx, ky)
        oR = lattice_open(energy, V_R, nx, ny, Lx, Ly, delta, kx, ky)
        return np.where(oL, sL.imag, 0.0) / delta ** 2, np.where(oR, sR.imag, 0.0) / delta ** 2
    kin = kinetic(nx, ny, Lx, Ly, kx, ky)
    QR = np.sqrt(np.maximum(2 * (energy - V_L) - kin, 0.0))
    QT = np.sqrt(np.maximum(2 * (energy - V_R) - kin, 0.0))
    return np.sin(QR * delta) / delta ** 2, np.sin(QT * delta) / delta ** 2


def transmission_solver(G_last, energy: float, nx, ny, Lx, Ly, Lz, N_z, kx, ky, V_L, V_R,
                        lead: str = "closed-form", eta: float = 0.0):
    """Transmission_solver (src/grid.f90:335-385) applied to first_column(:, :, N_z)
    (already rescaled by 2 Delta^2).  Real energy, no eta; closed channels give 0.
    tau = sum_ij gammaL_i gammaR_j |G(i, j)|^2 (the reference's L/R placement)."""
    gL, gR = lead_gammas(energy, nx, ny, Lx, Ly, Lz, N_z, kx, ky, V_L, V_R, lead, eta)
    A = gL[:, None] * G_last
    B = gR[:, None] * np.conj(G_last.T)
    s = np.sum(A * B.T)
    return float(s.real)


# ---------------------------------------------------------------- quadrature
def coefficient_simpson(i: int, N: int, a: float, b: float) -> float:
    """src/math_kernel.f90:167-204 (1-based i; N=2,4 quirks kept)."""
    if N % 2 == 1:
        if i == 1 or i == N:
            c = 1.0 / 3
        else:
            c = 4.0 / 3 if i % 2 == 0 else 2.0 / 3
    else:
        if i == 1:
            c = 1.0 / 3
        elif i == N - 3:
            c = 1.0 / 3 + 1.0 * 3 / 8
        elif i == N - 2 or i == N - 1:
            c = 3.0 * 3 / 8
        elif i == N:
            c = 1.0 * 3 / 8
        else:
            c = 4.0 / 3 if i % 2 == 0 else 2.0 / 3
    return c * (b - a) / (N - 1)


def value_simpson(i: int, N: int, a: float, b: float) -> float:
    """src/math_kernel.f90:206-216."""
    return (i - 1) * (b - a) / (N - 1) + a


def fermi_func(energy: float, mu: float, temperature: float) -> float:
    """src/math_kernel.f90:218-223."""
    return 1.0 / (math.exp((energy - mu) / (K_B * temperature)) + 1)


# ---------------------------------------------------------------- per-job drivers
@dataclass
class System:
    """Everything the per-job routines read: basis, V-hat table, parameters."""
    atomic: Atomic
    nx: np.ndarray
    ny: np.ndarray
    Vrec: np.ndarray          # (2NGX+1, 2NGY+1, N_z)
    NGX: int
    NGY: int
    kx: np.ndarray = field(default_factory=lambda: np.zeros(1))
    ky: np.ndarray = field(default_factory=lambda: np.zeros(1))
    kw: np.ndarray = field(default_factory=lambda: np.ones(1))
    lead: str = "closed-form"          # or "decimation" (extension)

    @property
    def N(self):
        return len(self.nx)

    @property
    def N_z(self):
        return self.Vrec.shape[2]


def _blocks(sys: System, ik: int, energy: complex):
    a = sys.atomic
    H = hamiltonian(sys.nx, sys.ny, a.LX, a.LY, sys.kx[ik], sys.ky[ik], sys.Vrec, sys.NGX, sys.NGY)
    return e_minus_h(H, energy, sys.nx, sys.ny, a.LX, a.LY, a.LZ, sys.kx[ik], sys.ky[ik],
                     a.V_L, a.V_R, sys.lead)


def transmission_energies(atomic: Atomic, N_E: int):
    """src/main.f90:369-375: last point excluded."""
    e0, e1 = atomic.TRANSMISSION_GRID
    return np.array([e0 + (e1 - e0) * (i - 1) / N_E for i in range(1, N_E + 1)])


def transmission_tau(sys: System, energy: float, ik: int, faithful_solver: bool = True):
    """Transmission_Coefficient (src/negf.f90:115-174) for one (energy, k) job;
    returns tau (before the k weight).  faithful_solver=False uses the O(N^2)-state
    restatement rgf_first_column_last_slice (same generators)."""
    a = sys.atomic
    D = _blocks(sys, ik, energy + 1j * a.ETA)
    if faithful_solver:
        G = tri_solver(D, "F")["first_column"][sys.N_z - 1]
    else:
        G = rgf_first_column_last_slice(D)
    G = G * 2.0 * (a.LZ / sys.N_z) ** 2
    return transmission_solver(G, energy, sys.nx, sys.ny, a.LX, a.LY, a.LZ, sys.N_z,
                               sys.kx[ik], sys.ky[ik], a.V_L, a.V_R, sys.lead, a.ETA)


def transmission(sys: System, energies, faithful_solver: bool = True):
    """The transmission job loop (src/main.f90:384-398) on one rank:
    T(E_i) = sum_k w_k tau(E_i, k)."""
    T = np.zeros(len(energies))
    for ik in range(len(sys.kx)):
        for ie, E in enumerate(energies):
            T[ie] += transmission_tau(sys, E, ik, faithful_solver) * sys.kw[ik]
    return T


def density_contour(atomic: Atomic, N_circle: int, N_line: int, min_V: float, quadrature: str = "simpson"):
    """Contour nodes and weights of Equilibrium_Density (src/negf.f90:36-65,86-111).
    Returns list of (energy (complex, without eta), weight (complex)) such that
    Density_Matrix += w_k * weight * G_kk (G_kk already rescaled by 2 Delta^2).
    quadrature="gauss-legendre" (extension, no reference counterpart) maps
    numpy.polynomial.legendre.leggauss onto the arc theta in [0, pi] and the real segment."""
    circle_L = min_V - 1.0 / HARTREE
    circle_R = atomic.MU - atomic.GAP
    line_L, line_R = atomic.MU - atomic.GAP, atomic.MU + atomic.GAP
    E_c, E_R = (circle_R + circle_L) / 2, (circle_R - circle_L) / 2
    nodes = []
    if quadrature == "gauss-legendre":
        xc, wc = np.polynomial.legendre.leggauss(N_circle) if N_circle else ([], [])
        for x, g in zip(xc, wc):
            theta = PI / 2 * (x + 1.0)
            nodes.append((complex(E_c + E_R * np.exp(1j * theta)),
                          complex(2.0 / PI * (PI / 2 * g) * 1j * E_R * np.exp(1j * theta))))
        xl, wl = np.polynomial.legendre.leggauss(N_line) if N_line else ([], [])
        half, mid = (line_R - line_L) / 2, (line_R + line_L) / 2
        for x, g in zip(xl, wl):
            E = mid + half * x
            nodes.append((complex(E), complex(-(2.0 / PI * half * g * fermi_func(E, atomic.MU, atomic.TEMPERATURE)))))
        return nodes
    for i in range(1, N_circle + N_line + 1):
        if i <= N_circle:
            theta = value_simpson(i, N_circle, 0.0, PI)
            E = E_c + E_R * np.exp(1j * theta)
            w = 2.0 / PI * coefficient_simpson(i, N_circle, 0.0, PI) * 1j * E_R * np.exp(1j * theta)
        else:
            E = complex(value_simpson(i - N_circle, N_line, line_L, line_R))
            w = -(2.0 / PI * coefficient_simpson(i - N_circle, N_line, line_L, line_R)
                  * fermi_func(E.real, atomic.MU, atomic.TEMPERATURE))
        nodes.append((complex(E), complex(w)))
    return nodes


def n_line(atomic: Atomic, N_line_per_eV: int) -> int:
    """src/main.f90:261: IDNINT(N_line_per_eV * 2 GAP[eV])."""
    x = N_line_per_eV * ((2 * atomic.GAP) * HARTREE)
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def diagonal_blocks(sys: System, energy: complex, ik: int):
    """tri_solver 'D' + 2 Delta^2 rescale (src/negf.f90:73-84): G_kk, (N_z, N, N)."""
    a = sys.atomic
    D = _blocks(sys, ik, energy + 1j * a.ETA)
    return tri_solver(D, "D")["diagonal"] * 2.0 * (a.LZ / sys.N_z) ** 2


def density_matrix(sys: System, N_circle: int, N_line: int, min_V: float, quadrature: str = "simpson"):
    """Equilibrium_Density summed over all jobs (src/negf.f90:9-113,
    src/main.f90:272-288): Density_Matrix(N, N, N_z) stored as (N_z, N, N)."""
    DM = np.zeros((sys.N_z, sys.N, sys.N), dtype=np.complex128)
    for ik in range(len(sys.kx)):
        for E, w in density_contour(sys.atomic, N_circle, N_line, min_V, quadrature):
            DM += sys.kw[ik] * w * diagonal_blocks(sys, E, ik)
    return DM


# ---------------------------------------------------------------- bias (extension)
# Non-equilibrium density for VDS != 0.  No reference counterpart: the reference rejects
# VDS != 0 (src/main.f90:129-132).  Convention: mu_L = MU, mu_R = MU - VDS.  The left
# lead is attached to slice 1 with gamma^L_i = sin(Q^L_i Delta) / Delta^2 (Q^L from V_L,
# real energy, closed channels 0 -- the trace's formula, src/grid.f90:355-360).
def gamma_lead(energy: float, V_lead: float, nx, ny, Lx, Ly, Lz, N_z, kx, ky):
    delta = Lz / N_z
    Q = np.sqrt(np.maximum(2 * (energy - V_lead) - kinetic(nx, ny, Lx, Ly, kx, ky), 0.0))
    return np.sin(Q * delta) / delta ** 2


def neq_spectral_blocks(sys: System, energy: float, ik: int, lead: str = "L"):
    """Diagonal blocks of A_L = G Gamma_L G^dagger (lead "L": Gamma on slice 1 from V_L)
    or A_R (lead "R": Gamma on slice N_z from V_R), physical units (G = 2 Delta^2 times
    the inverse of the scaled operator), by dense inversion of tridiag(I, D_k, I):
    (N_z, N, N)."""
    a = sys.atomic
    D = _blocks(sys, ik, energy + 1j * a.ETA)
    N_z, N = sys.N_z, sys.N
    G = np.linalg.inv(assemble_dense(D)) * 2.0 * (a.LZ / N_z) ** 2
    gL, gR = lead_gammas(energy, sys.nx, sys.ny, a.LX, a.LY, a.LZ, N_z, sys.kx[ik], sys.ky[ik],
                         a.V_L, a.V_R, sys.lead, a.ETA)
    gam = gL if lead == "L" else gR
    col = G[:, :N] if lead == "L" else G[:, (N_z - 1) * N:]     # G_{k,1} or G_{k,N_z}
    out = np.empty((N_z, N, N), dtype=np.complex128)
    for k in range(N_z):
        Gk1 = col[k * N:(k + 1) * N]
        out[k] = (Gk1 * gam[None, :]) @ Gk1.conj().T
    return out


def bias_window(atomic: Atomic, VDS: float, n_nodes: int, quadrature: str = "simpson"):
    """Real-axis nodes and weights of the bias-window integral
        Density_Matrix += i (1/pi) integral (f_L - f_R) A_L dE,  mu_L = MU, mu_R = MU - VDS
    over [min(mu) - GAP, max(mu) + GAP] (GAP in Hartree as t_parameters holds it).  The
    i makes Im(Density_Matrix), which the density reads, carry (1/pi) integral (f_L - f_R) A_L
    (2 for spin times 1/(2 pi)).  Returns (E, w) with w complex."""
    muL, muR = atomic.MU, atomic.MU - VDS
    lo, hi = min(muL, muR) - atomic.GAP, max(muL, muR) + atomic.GAP
    if quadrature == "gauss-legendre":
        x, g = np.polynomial.legendre.leggauss(n_nodes)
        E = (hi + lo) / 2 + (hi - lo) / 2 * x
        q = (hi - lo) / 2 * g
    else:
        E = np.array([value_simpson(i, n_nodes, lo, hi) for i in range(1, n_nodes + 1)])
        q = np.array([coefficient_simpson(i, n_nodes, lo, hi) for i in range(1, n_nodes + 1)])
    df = np.array([fermi_func(e, muL, atomic.TEMPERATURE) - fermi_func(e, muR, atomic.TEMPERATURE) for e in E])
    return E, 1j / PI * q * df


def density_matrix_bias(sys: System, N_circle: int, N_line: int, min_V: float, VDS: float, n_window: int,
                        quadrature: str = "simpson"):
    """Density_Matrix for VDS != 0: the equilibrium contour with the Fermi level at
    min(mu_L, mu_R), plus the bias window with A_L of the lead whose level is higher
    (VDS > 0: the left lead)."""
    a = sys.atomic
    mu_min = min(a.MU, a.MU - VDS)
    import dataclasses
    a_eq = dataclasses.replace(a, MU=mu_min)
    DM = np.zeros((sys.N_z, sys.N, sys.N), dtype=np.complex128)
    sys_eq = dataclasses.replace(sys, atomic=a_eq)
    for ik in range(len(sys.kx)):
        for E, w in density_contour(a_eq, N_circle, N_line, min_V, quadrature):
            DM += sys.kw[ik] * w * diagonal_blocks(sys_eq, E, ik)
    if VDS == 0.0:
        return DM
    if VDS < 0.0:
        raise ValueError("VDS < 0 (right lead higher) is not restated")
    Ew, ww = bias_window(a, VDS, n_window, quadrature)
    for ik in range(len(sys.kx)):
        for E, w in zip(Ew, ww):
            DM += sys.kw[ik] * w * neq_spectral_blocks(sys, float(E), ik)
    return DM


def ldos_energies(atomic: Atomic, N_E: int):
    """src/negf.f90:203-204: last point excluded."""
    e0, e1 = atomic.LDOS_ENERGY_GRID
    return np.array([e0 + (e1 - e0) * (i - 1) / N_E for i in range(1, N_E + 1)])


def ldos_matrix(sys: System, energy: float):
    """Local_Density_Of_State summed over k (src/negf.f90:176-236):
    LDOS_Matrix -= w_k / pi * G_kk."""
    LM = np.zeros((sys.N_z, sys.N, sys.N), dtype=np.complex128)
    for ik in range(len(sys.kx)):
        LM -= sys.kw[ik] / PI * diagonal_blocks(sys, energy, ik)
    return LM


# ---------------------------------------------------------------- plane wave -> real
def pw_to_real_literal(M: np.ndarray, nx, ny, N_x: int, N_y: int) -> np.ndarray:
    """Greensfunction_PlanewaveToReal (src/grid.f90:71-135), literally: build the
    (N_x, N_y, N_x, N_y) tensor, unnormalised FORWARD FFT over the second index pair,
    BACKWARD over the first, take the diagonal.  Tiny grids only."""
    T = np.zeros((N_x, N_y, N_x, N_y), dtype=np.complex128)
    for j in range(len(nx)):
        for i in range(len(nx)):
            T[(nx[i] + N_x // 2) - 1, (ny[i] + N_y // 2) - 1,
              (nx[j] + N_x // 2) - 1, (ny[j] + N_y // 2) - 1] = M[i, j]
    T = np.fft.fft2(T, axes=(2, 3))                         # FFTW_FORWARD
    T = np.fft.ifft2(T, axes=(0, 1)) * (N_x * N_y)          # FFTW_BACKWARD, unnormalised
    out = np.empty((N_x, N_y), dtype=np.complex128)
    for jj in range(N_y):
        for ii in range(N_x):
            out[ii, jj] = T[ii, jj, ii, jj]
    return out


def pw_to_real(M: np.ndarray, nx, ny, N_x: int, N_y: int) -> np.ndarray:
    """Same output as pw_to_real_literal in O(N^2 + N_x N_y log): the diagonal of the
    double transform is out(m) = sum_ij M_ij exp(+2 pi i (dn_x m_x/N_x + dn_y m_y/N_y))
    with dn = n_i - n_j (the index offsets of :92-93 cancel on the diagonal).
    rho(Q) = sum_{n_i - n_j = Q mod grid} M_ij, then an unnormalised inverse DFT."""
    rho = np.zeros((N_x, N_y), dtype=np.complex128)
    qx = (nx[:, None] - nx[None, :]) % N_x
    qy = (ny[:, None] - ny[None, :]) % N_y
    np.add.at(rho, (qx.ravel(), qy.ravel()), M.ravel())
    return np.fft.ifft2(rho) * (N_x * N_y)


def density_real(DM: np.ndarray, nx, ny, N_x, N_y, LX_ang, LY_ang, LZ_ang):
    """src/main.f90:326-350: per slice transform, Density = Im(.) * N_z / (LX LY LZ)
    in Angstrom^-3.  Returns (N_x, N_y, N_z)."""
    N_z = DM.shape[0]
    out = np.empty((N_x, N_y, N_z))
    r = N_z / (LX_ang * LY_ang * LZ_ang)
    for k in range(N_z):
        out[:, :, k] = pw_to_real(DM[k], nx, ny, N_x, N_y).imag * r
    return out


def ldos_size(v: int, n: int) -> int:
    """src/tools.f90:6-16."""
    return n if v == 0 else 1


def ldos_grid_converter(i: int, v: int, delta_i: float, rescale: float):
    """src/tools.f90:18-40 (1-based i).  Returns (ii, rescale)."""
    ii = i if v == 0 else 1
    if v == 0:
        pass
    elif v == -1:
        rescale = rescale * delta_i
    elif v == i:
        pass
    else:
        rescale = 0.0
    return ii, rescale


def ldos_reduce(LM: np.ndarray, nx, ny, N_x, N_y, LX_ang, LY_ang, LZ_ang, LDOS_GRID):
    """src/main.f90:481-525 for one energy: transform every slice, then select /
    integrate / keep per dimension and rescale to Angstrom^-n eV^-1."""
    N_z = LM.shape[0]
    mx, my, mz = (ldos_size(LDOS_GRID[0], N_x), ldos_size(LDOS_GRID[1], N_y),
                  ldos_size(LDOS_GRID[2], N_z))
    out = np.zeros((mx, my, mz))
    for k in range(1, N_z + 1):
        im = pw_to_real(LM[k - 1], nx, ny, N_x, N_y).imag
        for j in range(1, N_y + 1):
            for i in range(1, N_x + 1):
                r = N_z / (LX_ang * LY_ang * LZ_ang) / HARTREE
                ii, r = ldos_grid_converter(i, LDOS_GRID[0], LX_ang / N_x, r)
                jj, r = ldos_grid_converter(j, LDOS_GRID[1], LY_ang / N_y, r)
                kk, r = ldos_grid_converter(k, LDOS_GRID[2], LZ_ang / N_z, r)
                out[ii - 1, jj - 1, kk - 1] += im[i - 1, j - 1] * r
    return out


# ---------------------------------------------------------------- brute force (pins)
def assemble_dense(G_inv_D: np.ndarray) -> np.ndarray:
    """The full (N N_z)^2 block-tridiagonal matrix tridiag(I, D_k, I) (Appendix of the
    RGF note referenced at src/math_kernel.f90:46-50).  Used only by pin tests."""
    N_z, N, _ = G_inv_D.shape
    A = np.zeros((N * N_z, N * N_z), dtype=np.complex128)
    I = np.eye(N)
    for k in range(N_z):
        A[k * N:(k + 1) * N, k * N:(k + 1) * N] = G_inv_D[k]
        if k + 1 < N_z:
            A[k * N:(k + 1) * N, (k + 1) * N:(k + 2) * N] = I
            A[(k + 1) * N:(k + 2) * N, k * N:(k + 1) * N] = I
    return A
