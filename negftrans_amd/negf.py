"""Python host side of the MI355X RGF energy loop, mirroring NEGFtrans's interfaces.

Every numeric routine of the energy loop runs in libnegf_gpu.so (HIP, gfx950) through
the C-ABI of include/negf_gpu.h; this module only prepares arguments the way the
reference's host code does (basis, k mesh, energy grids, contour nodes, unit
conversion) and returns results in the reference's array layouts (Fortran order,
(N, N, N_z) blocks).  Citations are to the NEGFtrans Fortran sources.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr

# src/global.f90:24-28
HARTREE = 27.211386245988
K_B = 8.617333262e-5 / HARTREE
A_0 = 0.529177210903
PI = 3.141592653589793

MODE_D, MODE_F, MODE_L = 1, 2, 4
PREC_FP64, PREC_MIXED_FP32 = 0, 1
FULL_WINDOW = 16
LEAD_DECIMATION = 32     # extension: lattice lead terms by decimation on the GPU
AUTOTUNE = 64            # extension: FP64 transmission kernel shape and form picked by timing


# ------------------------------------------------------------------ parameters
@dataclass
class Atomic:
    """t_parameters (src/global.f90:5-7) in atomic units."""
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

    @classmethod
    def from_input(cls, V_L=0.0, V_R=0.0, MU=5.568, ETA=1e-5, TEMPERATURE=20.0, LX=0.0, LY=0.0,
                   LZ=0.0, ENCUT=100.0, GAP=6.0, VDS=0.0, TRANSMISSION_GRID=(0.0, 0.0, 0.0),
                   LDOS_ENERGY_GRID=(0.0, 0.0, 0.0), **_ignored):
        """INPUT namelist values (eV, Angstrom, K; ETA in Hartree) -> atomic units
        (src/main.f90:182-198)."""
        return cls(V_L=V_L / HARTREE, V_R=V_R / HARTREE, MU=MU / HARTREE, ETA=ETA,
                   TEMPERATURE=TEMPERATURE, LX=LX / A_0, LY=LY / A_0, LZ=LZ / A_0,
                   ENCUT=ENCUT / HARTREE, GAP=GAP * K_B * TEMPERATURE, VDS=VDS / HARTREE,
                   TRANSMISSION_GRID=(TRANSMISSION_GRID[0] / HARTREE, TRANSMISSION_GRID[1] / HARTREE),
                   LDOS_ENERGY_GRID=(LDOS_ENERGY_GRID[0] / HARTREE, LDOS_ENERGY_GRID[1] / HARTREE))


def ngx_auto(NGX, NGY, LX_ang, LY_ang, ENCUT_eV):
    """src/main.f90:94-100 (the NGX test compares against the LY bound, as there)."""
    bx = math.ceil((LX_ang / A_0) / PI * math.sqrt(ENCUT_eV / HARTREE / 2) * 2)
    by = math.ceil((LY_ang / A_0) / PI * math.sqrt(ENCUT_eV / HARTREE / 2) * 2)
    if NGX <= by:
        NGX = bx + 1
    if NGY <= by:
        NGY = by + 1
    return NGX, NGY


def planewave_basis(ENCUT, Lx, Ly):
    """PlaneWaveBasis_construction (src/grid.f90:387-431): n_y outer, n_x inner."""
    nxm = math.ceil(Lx / PI * math.sqrt(ENCUT / 2))
    nym = math.ceil(Ly / PI * math.sqrt(ENCUT / 2))
    j, i = np.meshgrid(np.arange(-nym, nym + 1), np.arange(-nxm, nxm + 1), indexing="ij")
    i, j = i.ravel(), j.ravel()
    keep = ((2 * i * PI / Lx) ** 2 + (2 * j * PI / Ly) ** 2) / 2 <= ENCUT
    return i[keep].astype(np.int32), j[keep].astype(np.int32)


def kpoint_mesh(NKX, NKY, Lx, Ly):
    """Kpoint_mesh_construction (src/grid.f90:433-448): x fastest."""
    i = np.tile(np.arange(1, NKX + 1), NKY)
    j = np.repeat(np.arange(1, NKY + 1), NKX)
    kx = (i - (1 + NKX) / 2.0) / NKX * (2 * PI / Lx)
    ky = (j - (1 + NKY) / 2.0) / NKY * (2 * PI / Ly)
    w = np.full(NKX * NKY, 1.0 / (NKX * NKY))
    return kx, ky, w


def uniform_energies(e0, e1, n):
    """src/main.f90:369-375 / src/negf.f90:203-204: E_i = e0 + (e1-e0)(i-1)/n, i=1..n."""
    return np.array([e0 + (e1 - e0) * (i - 1) / n for i in range(1, n + 1)])


def coefficient_simpson(i, N, a, b):
    """src/math_kernel.f90:167-204 (1-based i)."""
    if N % 2 == 1:
        c = 1.0 / 3 if i in (1, N) else (4.0 / 3 if i % 2 == 0 else 2.0 / 3)
    elif i == 1:
        c = 1.0 / 3
    elif i == N - 3:
        c = 1.0 / 3 + 1.0 * 3 / 8
    elif i in (N - 2, N - 1):
        c = 3.0 * 3 / 8
    elif i == N:
        c = 1.0 * 3 / 8
    else:
        c = 4.0 / 3 if i % 2 == 0 else 2.0 / 3
    return c * (b - a) / (N - 1)


def value_simpson(i, N, a, b):
    """src/math_kernel.f90:206-216."""
    return (i - 1) * (b - a) / (N - 1) + a


def fermi_func(energy, mu, temperature):
    """src/math_kernel.f90:218-223."""
    return 1.0 / (math.exp((energy - mu) / (K_B * temperature)) + 1)


def n_line(atomic: Atomic, N_line_per_eV: int) -> int:
    """src/main.f90:261: IDNINT(N_line_per_eV * 2 GAP[eV])."""
    x = N_line_per_eV * ((2 * atomic.GAP) * HARTREE)
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def gauss_legendre(n: int):
    """Nodes and weights of the n-point Gauss-Legendre rule on [-1, 1]: Newton iteration
    on P_n from the three-term recurrence, started at the asymptotic root estimate
    cos(pi (i - 1/4) / (n + 1/2)); w_i = 2 / ((1 - x_i^2) P_n'(x_i)^2)."""
    if n < 1:
        raise ValueError("n must be >= 1")
    x = np.cos(np.pi * (np.arange(1, n + 1) - 0.25) / (n + 0.5))
    for _ in range(100):
        p0, p1 = np.ones_like(x), x.copy()
        for k in range(2, n + 1):
            p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
        dp = n * (x * p1 - p0) / (x * x - 1.0)
        dx = p1 / dp
        x = x - dx
        if np.max(np.abs(dx)) < 1e-15:
            break
    p0, p1 = np.ones_like(x), x.copy()
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    dp = n * (x * p1 - p0) / (x * x - 1.0)
    w = 2.0 / ((1.0 - x * x) * dp * dp)
    order = np.argsort(x)
    return x[order], w[order]


def density_contour(atomic: Atomic, N_circle: int, N_line: int, min_V: float, quadrature: str = "simpson"):
    """Contour nodes z_n (without eta) and weights w_n of Equilibrium_Density
    (src/negf.f90:36-65, 86-111); Density_Matrix += w_k w_n G_kk(z_n + i eta).

    quadrature = "simpson" is the reference's rule (coefficient_simpson / value_simpson,
    src/math_kernel.f90:167-204).  "gauss-legendre" (an extension) puts N_circle
    Gauss-Legendre nodes on the arc theta in [0, pi] and N_line on the real segment, with
    the same integrands."""
    circle_L = min_V - 1.0 / HARTREE
    circle_R = atomic.MU - atomic.GAP
    line_L, line_R = atomic.MU - atomic.GAP, atomic.MU + atomic.GAP
    E_c, E_R = (circle_R + circle_L) / 2, (circle_R - circle_L) / 2
    z = np.empty(N_circle + N_line, dtype=np.complex128)
    w = np.empty_like(z)
    if quadrature == "simpson":
        for i in range(1, N_circle + N_line + 1):
            if i <= N_circle:
                th = value_simpson(i, N_circle, 0.0, PI)
                e = np.exp(1j * th)
                z[i - 1] = E_c + E_R * e
                w[i - 1] = 2.0 / PI * coefficient_simpson(i, N_circle, 0.0, PI) * 1j * E_R * e
            else:
                E = value_simpson(i - N_circle, N_line, line_L, line_R)
                z[i - 1] = E
                w[i - 1] = -(2.0 / PI * coefficient_simpson(i - N_circle, N_line, line_L, line_R)
                             * fermi_func(E, atomic.MU, atomic.TEMPERATURE))
        return z, w
    if quadrature != "gauss-legendre":
        raise ValueError(f"unknown contour quadrature {quadrature!r}")
    if N_circle > 0:
        x, g = gauss_legendre(N_circle)
        th = PI / 2 * (x + 1.0)
        e = np.exp(1j * th)
        z[:N_circle] = E_c + E_R * e
        w[:N_circle] = 2.0 / PI * (PI / 2 * g) * 1j * E_R * e
    if N_line > 0:
        x, g = gauss_legendre(N_line)
        half, mid = (line_R - line_L) / 2, (line_R + line_L) / 2
        E = mid + half * x
        z[N_circle:] = E
        w[N_circle:] = -(2.0 / PI * half * g * np.array([fermi_func(e_, atomic.MU, atomic.TEMPERATURE) for e_ in E]))
    return z, w


def bias_window(atomic: Atomic, VDS: float, n_nodes: int, quadrature: str = "simpson"):
    """Nodes E_n and weights w_n of the bias-window integral of the non-equilibrium density
    (extension; the reference rejects VDS != 0, src/main.f90:129-132):
        Density_Matrix += i (1/pi) integral (f_L - f_R) [G Gamma_L G^dagger] dE
    with mu_L = MU, mu_R = MU - VDS (VDS > 0) over [mu_R - GAP, mu_L + GAP].  The i puts
    the term into Im(Density_Matrix), which the density reads."""
    if VDS <= 0.0:
        raise ValueError("bias_window needs VDS > 0 (mu_L = MU above mu_R = MU - VDS)")
    muL, muR = atomic.MU, atomic.MU - VDS
    lo, hi = muR - atomic.GAP, muL + atomic.GAP
    if quadrature == "gauss-legendre":
        x, g = gauss_legendre(n_nodes)
        E = (hi + lo) / 2 + (hi - lo) / 2 * x
        q = (hi - lo) / 2 * g
    elif quadrature == "simpson":
        E = np.array([value_simpson(i, n_nodes, lo, hi) for i in range(1, n_nodes + 1)])
        q = np.array([coefficient_simpson(i, n_nodes, lo, hi) for i in range(1, n_nodes + 1)])
    else:
        raise ValueError(f"unknown quadrature {quadrature!r}")
    df = np.array([fermi_func(e, muL, atomic.TEMPERATURE) - fermi_func(e, muR, atomic.TEMPERATURE) for e in E])
    return E, 1j / PI * q * df


# ------------------------------------------------------------------ GPU entry points
def device_count():
    return _lib.load().negf_device_count()


def local_potential_real_to_planewave(V_real, NGX, NGY, full_window=False):
    """V_reciprocal_all(-NGX:NGX, -NGY:NGY, N_z) from V_real(N_x, N_y, N_z) (Hartree):
    LocalPotential_RealToPlanewave per slice + the window copy (src/grid.f90:12-69,
    src/main.f90:238-247).  Returned as a Fortran-ordered (2NGX+1, 2NGY+1, N_z) array."""
    V = np.asfortranarray(np.asarray(V_real, dtype=np.float64))
    N_x, N_y, N_z = V.shape
    out = np.zeros((2 * NGX + 1, 2 * NGY + 1, N_z), dtype=np.complex128, order="F")
    lib = _lib.load()
    check(lib.negf_vhat_from_real(N_x, N_y, N_z, dptr(V), NGX, NGY, int(bool(full_window)),
                                  dptr(out)), "negf_vhat_from_real")
    return out


def nonlocal_potential_real_to_planewave(V_real, Rx, Ry):
    """NonLocalPotential_RealToPlanewave (src/grid.f90:137-234) on the GPU:
    V_real(N_x, N_y, N_x, N_y) complex -> V_reciprocal(-Rx:Rx, -Ry:Ry, -Rx:Rx, -Ry:Ry),
    returned as a Fortran-ordered array indexed [p + Rx, q + Ry, p' + Rx, q' + Ry]."""
    V = np.asfortranarray(np.asarray(V_real, dtype=np.complex128))
    N_x, N_y = V.shape[:2]
    if V.shape != (N_x, N_y, N_x, N_y):
        raise ValueError("V_real must be (N_x, N_y, N_x, N_y)")
    out = np.zeros((2 * Rx + 1, 2 * Ry + 1, 2 * Rx + 1, 2 * Ry + 1), dtype=np.complex128, order="F")
    check(_lib.load().negf_nonlocal_vhat_from_real(N_x, N_y, dptr(V), Rx, Ry, dptr(out)),
          "negf_nonlocal_vhat_from_real")
    return out


def inverse(M):
    """Batched drop-in for inverse (src/math_kernel.f90:8-34). M: (N, N) or (N, N, B)."""
    A = np.asfortranarray(np.asarray(M, dtype=np.complex128))
    squeeze = A.ndim == 2
    if squeeze:
        A = A[:, :, None]
    N, _, B = A.shape
    out = np.zeros_like(A, order="F")
    check(_lib.load().negf_inverse_batch(N, B, dptr(A), dptr(out)), "negf_inverse_batch")
    return out[:, :, 0] if squeeze else out


def matrix_product(A, B):
    """Batched drop-in for matrix_product (src/math_kernel.f90:36-42)."""
    A = np.asfortranarray(np.asarray(A, dtype=np.complex128))
    Bm = np.asfortranarray(np.asarray(B, dtype=np.complex128))
    squeeze = A.ndim == 2
    if squeeze:
        A, Bm = A[:, :, None], Bm[:, :, None]
    N, _, nb = A.shape
    C = np.zeros_like(A, order="F")
    check(_lib.load().negf_matrix_product_batch(N, nb, dptr(A), dptr(Bm), dptr(C)),
          "negf_matrix_product_batch")
    return C[:, :, 0] if squeeze else C


def greensfunction_tri_solver(G_inv_D, DFL):
    """GreensFunction_tri_solver (src/math_kernel.f90:44-165).  G_inv_D: (N, N, N_z)
    complex128; returns dict of (N, N, N_z) arrays 'diagonal', 'first_column',
    'last_column' (zero where not requested) and the status code."""
    D = np.asfortranarray(np.asarray(G_inv_D, dtype=np.complex128))
    N, _, N_z = D.shape
    mode = (MODE_D if any(c in DFL for c in "Dd") else 0) | \
           (MODE_F if any(c in DFL for c in "Ff") else 0) | \
           (MODE_L if any(c in DFL for c in "Ll") else 0)
    outs = {k: np.zeros((N, N, N_z), dtype=np.complex128, order="F")
            for k in ("diagonal", "first_column", "last_column")}
    nul = ctypes.POINTER(ctypes.c_double)()
    st = check(_lib.load().negf_rgf_tri_solve(
        N, N_z, dptr(D), mode,
        dptr(outs["diagonal"]) if mode & MODE_D else nul,
        dptr(outs["first_column"]) if mode & MODE_F else nul,
        dptr(outs["last_column"]) if mode & MODE_L else nul), "negf_rgf_tri_solve")
    outs["status"] = st
    return outs


def greensfunction_planewave_to_real(M, nx, ny, N_x, N_y):
    """Greensfunction_PlanewaveToReal (src/grid.f90:71-135) for one (N, N) block or a
    stack (N, N, B); returns (N_x, N_y) or (N_x, N_y, B) complex."""
    A = np.asfortranarray(np.asarray(M, dtype=np.complex128))
    squeeze = A.ndim == 2
    if squeeze:
        A = A[:, :, None]
    N, _, B = A.shape
    nx = np.ascontiguousarray(nx, dtype=np.int32)
    ny = np.ascontiguousarray(ny, dtype=np.int32)
    out = np.zeros((N_x, N_y, B), dtype=np.complex128, order="F")
    check(_lib.load().negf_pw_to_real(N, iptr(nx), iptr(ny), B, dptr(A), N_x, N_y, dptr(out)),
          "negf_pw_to_real")
    return out[:, :, 0] if squeeze else out


class NegfDevice:
    """One transport system resident on the GPU (basis + V_reciprocal_all + geometry).

    Replaces the per-job setup of src/negf.f90 (Hamiltonian_construction and
    E_minus_H_construction, rebuilt for every energy there) with blocks generated on
    the fly inside the sweep kernels."""

    def __init__(self, nx, ny, atomic: Atomic, vrec, NGX, NGY, flags=PREC_FP64):
        self.nx = np.ascontiguousarray(nx, dtype=np.int32)
        self.ny = np.ascontiguousarray(ny, dtype=np.int32)
        self.atomic = atomic
        self.vrec = np.asfortranarray(np.asarray(vrec, dtype=np.complex128))
        self.NGX, self.NGY = NGX, NGY
        self.N = len(self.nx)
        self.N_z = self.vrec.shape[2]
        assert self.vrec.shape[:2] == (2 * NGX + 1, 2 * NGY + 1)
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        check(self._lib.negf_ctx_create(ctypes.byref(h), self.N, self.N_z, iptr(self.nx),
                                        iptr(self.ny), atomic.LX, atomic.LY, atomic.LZ, atomic.V_L,
                                        atomic.V_R, atomic.ETA, dptr(self.vrec), NGX, NGY, flags),
              "negf_ctx_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.negf_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def last_kernel_seconds(self):
        return self._lib.negf_last_kernel_seconds(self._h)

    @property
    def last_fallbacks(self):
        """Energies of the last mixed-precision transmission call recomputed in FP64."""
        return self._lib.negf_last_fallbacks(self._h)

    @property
    def last_variant(self):
        """Kernel variant of the last FP64 transmission call (0 one-sided two-tile, 1
        two-sided two-tile, 2 one-sided four-tile, 3 two-sided four-tile)."""
        return self._lib.negf_last_variant(self._h)

    def tune_seconds(self):
        """Device seconds of each variant in the context's last tuning call (AUTOTUNE),
        None for a variant that was not run."""
        out = [self._lib.negf_tune_seconds(self._h, v) for v in range(4)]
        return [x if x >= 0 else None for x in out]

    def set_potential(self, vrec):
        """Replace V_reciprocal_all (same window and N_z), e.g. in a self-consistent loop."""
        v = np.asfortranarray(np.asarray(vrec, dtype=np.complex128))
        if v.shape != self.vrec.shape:
            raise ValueError(f"V_reciprocal_all shape {v.shape} != {self.vrec.shape}")
        self.vrec = v
        check(self._lib.negf_ctx_set_potential(self._h, dptr(self.vrec)), "negf_ctx_set_potential")

    @staticmethod
    def _kp(kpoints):
        if kpoints is None:
            kpoints = (np.zeros(1), np.zeros(1), np.ones(1))
        kx, ky, kw = (np.ascontiguousarray(a, dtype=np.float64) for a in kpoints)
        return kx, ky, kw

    def transmission(self, energies, kpoints=None, per_job=False):
        """T(E_i) = sum_k w_k tau(E_i, k) for real energies (Hartree): the transmission
        job loop (src/main.f90:384-398) over Transmission_Coefficient
        (src/negf.f90:115-174), all jobs in one launch."""
        E = np.ascontiguousarray(energies, dtype=np.float64)
        kx, ky, kw = self._kp(kpoints)
        tau = np.zeros(len(E))
        jobs = np.zeros(len(E) * len(kx))
        st = check(self._lib.negf_transmission_batch(self._h, dptr(E), len(E), dptr(kx), dptr(ky),
                                                     dptr(kw), len(kx), dptr(tau), dptr(jobs)),
                   "negf_transmission_batch")
        self.last_status = st
        return (tau, jobs.reshape(len(kx), len(E))) if per_job else tau

    def neq_density_matrix(self, E, w, kpoints=None, out=None):
        """Bias-window part of the non-equilibrium density (extension):
        Density_Matrix += sum_k w_k sum_n w_n [G Gamma_L G^dagger]_kk(E_n + i eta), with the
        nodes and weights of bias_window."""
        EE = np.ascontiguousarray(E, dtype=np.float64)
        ww = np.ascontiguousarray(w, dtype=np.complex128)
        kx, ky, kw = self._kp(kpoints)
        if out is None:
            out = np.zeros((self.N, self.N, self.N_z), dtype=np.complex128, order="F")
        self.last_status = check(self._lib.negf_neq_density_batch(
            self._h, dptr(EE), dptr(ww), len(EE), dptr(kx), dptr(ky), dptr(kw), len(kx), dptr(out)),
            "negf_neq_density_batch")
        return out

    def density_matrix(self, z, w, kpoints=None, out=None):
        """Density_Matrix(N, N, N_z) += sum_k sum_n w_k w_n G_kk(z_n + i eta)
        (src/main.f90:272-288 over Equilibrium_Density, src/negf.f90:9-113)."""
        zz = np.ascontiguousarray(z, dtype=np.complex128)
        ww = np.ascontiguousarray(w, dtype=np.complex128)
        kx, ky, kw = self._kp(kpoints)
        if out is None:
            out = np.zeros((self.N, self.N, self.N_z), dtype=np.complex128, order="F")
        self.last_status = check(self._lib.negf_density_batch(
            self._h, dptr(zz), dptr(ww), len(zz), dptr(kx), dptr(ky), dptr(kw), len(kx), dptr(out)),
            "negf_density_batch")
        return out

    def ldos_matrix(self, energies, kpoints=None):
        """LDOS_Matrix(N, N, N_z, nE) = -sum_k w_k/pi G_kk(E + i eta)
        (Local_Density_Of_State, src/negf.f90:176-236)."""
        E = np.ascontiguousarray(energies, dtype=np.float64)
        kx, ky, kw = self._kp(kpoints)
        out = np.zeros((self.N, self.N, self.N_z, len(E)), dtype=np.complex128, order="F")
        self.last_status = check(self._lib.negf_ldos_batch(
            self._h, dptr(E), len(E), dptr(kx), dptr(ky), dptr(kw), len(kx), dptr(out)),
            "negf_ldos_batch")
        return out


# ------------------------------------------------------------------ output reductions
def density_from_matrix(DM, nx, ny, N_x, N_y, LX_ang, LY_ang, LZ_ang):
    """src/main.f90:306-350: plane-wave -> real space per slice on the GPU, then
    Density = Im(.) * N_z / (LX LY LZ) [Angstrom^-3]; returns (N_x, N_y, N_z)."""
    N_z = DM.shape[2]
    out = greensfunction_planewave_to_real(DM, nx, ny, N_x, N_y)
    return out.imag * (N_z / (LX_ang * LY_ang * LZ_ang))


def ldos_size(v, n):
    """src/tools.f90:6-16."""
    return n if v == 0 else 1


def ldos_reduce(LM, nx, ny, N_x, N_y, LX_ang, LY_ang, LZ_ang, LDOS_GRID):
    """src/main.f90:481-525 + LDOS_grid_converter (src/tools.f90:18-40) for one energy:
    transform every slice, keep (0) / integrate (-1) / select (m) per dimension,
    rescale to Angstrom^-n eV^-1."""
    N_z = LM.shape[2]
    im = greensfunction_planewave_to_real(LM, nx, ny, N_x, N_y).imag   # (N_x, N_y, N_z)
    base = N_z / (LX_ang * LY_ang * LZ_ang) / HARTREE
    axes = []
    for v, n, d in ((LDOS_GRID[0], N_x, LX_ang / N_x), (LDOS_GRID[1], N_y, LY_ang / N_y),
                    (LDOS_GRID[2], N_z, LZ_ang / N_z)):
        if v == 0:
            axes.append(("keep", None, 1.0))
        elif v == -1:
            axes.append(("sum", None, d))
        else:
            axes.append(("pick", v, 1.0))
    out = im * base
    for ax, (kind, m, f) in enumerate(axes):
        if kind == "sum":
            out = out.sum(axis=ax, keepdims=True) * f
        elif kind == "pick":
            n = out.shape[ax]
            sel = np.zeros(n)
            if 1 <= m <= n:
                sel[m - 1] = 1.0
            shape = [1, 1, 1]
            shape[ax] = n
            out = (out * sel.reshape(shape)).sum(axis=ax, keepdims=True)
    return out


class Multi:
    """Uneven batch (extension): several NegfDevice systems of one GPU, any N and N_z,
    computed in one hipGraph by negf_multi_transmission.  Each device's result is that of
    its own NegfDevice.transmission."""

    def __init__(self, devices):
        self.devices = list(devices)
        self._lib = _lib.load()
        hs = (ctypes.c_void_p * len(self.devices))(*[d._h.value for d in self.devices])
        h = ctypes.c_void_p()
        check(self._lib.negf_multi_create(ctypes.byref(h), len(self.devices), hs), "negf_multi_create")
        self._h = h
        self._tau = None

    def close(self):
        if getattr(self, "_h", None):
            self._lib.negf_multi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ptrs(self, arrays):
        return (ctypes.POINTER(ctypes.c_double) * len(arrays))(*[dptr(a) for a in arrays])

    def transmission(self, energies, kpoints=None):
        """energies: one array per device (Hartree); kpoints: None (Gamma only) or one
        (kx, ky, kw) per device.  Returns one T(E) array per device."""
        n = len(self.devices)
        if len(energies) != n:
            raise ValueError("one energy array per device")
        self._E = [np.ascontiguousarray(e, dtype=np.float64) for e in energies]
        kps = [NegfDevice._kp(None if kpoints is None else kpoints[d]) for d in range(n)]
        self._kx, self._ky, self._kw = ([k[i] for k in kps] for i in range(3))
        nE = np.array([len(e) for e in self._E], dtype=np.int32)
        nk = np.array([len(k) for k in self._kx], dtype=np.int32)
        self._tau = [np.zeros(len(e)) for e in self._E]
        self.last_status = check(self._lib.negf_multi_transmission(
            self._h, self._ptrs(self._E), iptr(nE), self._ptrs(self._kx), self._ptrs(self._ky),
            self._ptrs(self._kw), iptr(nk), self._ptrs(self._tau)), "negf_multi_transmission")
        return [t.copy() for t in self._tau]

    def ldos(self, energies, slices, kpoints=None):
        """LDOS blocks of every device at the given slices (0-based, the same number per
        device): one complex array (N, N, len(slices[d]), len(energies[d])) per device,
        equal to NegfDevice.ldos_matrix restricted to those slices."""
        n = len(self.devices)
        E = [np.ascontiguousarray(e, dtype=np.float64) for e in energies]
        kps = [NegfDevice._kp(None if kpoints is None else kpoints[d]) for d in range(n)]
        kx, ky, kw = ([k[i] for k in kps] for i in range(3))
        sl = [np.ascontiguousarray(s, dtype=np.int32) for s in slices]
        nsel = len(sl[0])
        if any(len(s) != nsel for s in sl):
            raise ValueError("the same number of slices per device")
        out = [np.zeros((d.N, d.N, nsel, len(e)), dtype=np.complex128, order="F") for d, e in zip(self.devices, E)]
        nE = np.array([len(e) for e in E], dtype=np.int32)
        nk = np.array([len(k) for k in kx], dtype=np.int32)
        sptr = (ctypes.POINTER(ctypes.c_int) * n)(*[iptr(s) for s in sl])
        self.last_status = check(self._lib.negf_multi_ldos(
            self._h, self._ptrs(E), iptr(nE), self._ptrs(kx), self._ptrs(ky), self._ptrs(kw), iptr(nk), nsel, sptr,
            self._ptrs(out)), "negf_multi_ldos")
        return out

    def replay(self):
        """Launch the last call's hipGraph again (after set_potential on any device)."""
        if self._tau is None:
            raise RuntimeError("replay needs a previous transmission call")
        tau = [np.zeros_like(t) for t in self._tau]
        self.last_status = check(self._lib.negf_multi_replay(self._h, self._ptrs(tau)), "negf_multi_replay")
        return tau

    @property
    def last_seconds(self):
        return self._lib.negf_multi_last_seconds(self._h)

    @staticmethod
    def variant(Np):
        """Inversion variant chosen for padded size Np: 'mfma', 'valu' or None."""
        v = _lib.load().negf_multi_variant(int(Np))
        return {0: "mfma", 1: "valu"}.get(v)

