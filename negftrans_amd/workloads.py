"""Synthetic transport systems of the benchmark and parity configurations.

Inputs are produced exactly as a NEGFtrans user would write INPUT and POTENTIAL
(eV, Angstrom; ETA in Hartree), then converted the reference's way.  The effective
mass m* enters through the exact scaling identity m* H*[V] = H_ref[m* V]: a GaAs
system at energy E is the reference (m = m_e) system at m* E with potential m* V and
broadening m* eta, so T*(E) = T_ref(m* E).  The plane-wave basis is fixed by its
size N (the cutoff radius R^2 in units of (2 pi / L)^2), not by m*.

Configurations (all single-device, k = 0):
  qpc_c0    Lx = Ly = 40 nm, N = 61 (R^2 = 19), N_z = 200 slices of 0.5 nm, GaAs,
            split-gate saddle V0 exp(-(z-50nm)^2/(2 (10nm)^2)) (x-Lx/2)^2/w^2,
            V0 = 10 meV, w = 8 nm, plus a 2 meV parabolic confinement in y
            (1/2 m* w^2 (y-Ly/2)^2 with hbar w = 2 meV); 400 energies in 0-20 meV,
            eta = 1e-6 eV.
  disorder_c1  Lx = Ly = 40 nm, N = 317 (R^2 = 100.5), N_z = 1000 slices of 0.25 nm,
            64 x 64 x 1000 grid of Gaussian-correlated disorder (sigma = 3 meV,
            correlation length 5 nm, periodic in x, y, seed 20261203), tapered to zero
            over 10 nm at both z ends; V_L = V_R = 0; 4096 energies in 0-30 meV.
  disorder_c3  disorder_c1's kind with N = 797 (R^2 = 256.5), N_z = 2000, sigma = 5 meV,
            80 x 80 x 2000 grid, 1024 energies in 0-30 meV.
  uneven_c2()  48 devices, N in {61, 113, 197, 317, 509}, N_z in 200-1500, seeded
            disorder of disorder_c1's kind each, 256 energies (the uneven batch).
  qpc_c4_instance(seed)  the learned adaptive grid's instances: qpc_c0's form with
            N = 113 (R^2 = 36.5), N_z = 400 slices of 0.25 nm on a 32 x 32 grid, seeded
            V0 in 6-14 meV, w in 6-10 nm, saddle centre 40-60 nm and width 8-12 nm,
            confinement 1.5-2.5 meV, plus Gaussian-correlated disorder (sigma drawn from
            0-0.5 meV by default, or the range `sigma` in eV; 5 nm); energies in 0-20 meV.
            From about 2 meV of disorder on, T(E) has resonances that uniform grids
            resolve only with thousands of points.  Features: the 16 x 16 x 8 lowest Fourier
            components of V, returned as the band-limited real samples they define.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .negf import A_0, HARTREE, Atomic, ngx_auto, planewave_basis, uniform_energies

MSTAR_GAAS = 0.067
HBAR2_OVER_ME = 7.619964   # hbar^2/m_e in eV Angstrom^2 (CODATA: 2 * 3.80998 eV A^2)


@dataclass
class Workload:
    name: str
    inp: dict                  # INPUT namelist values (eV, Angstrom, K; ETA in Hartree)
    V_real_eV: np.ndarray      # POTENTIAL (N_x, N_y, N_z) in eV (already scaled by m*)
    energies_eV: np.ndarray    # physical energies (eV); the reference sees m* E
    mstar: float = MSTAR_GAAS
    extra: dict = field(default_factory=dict)

    def atomic(self) -> Atomic:
        return Atomic.from_input(**self.inp)

    def basis(self):
        a = self.atomic()
        return planewave_basis(a.ENCUT, a.LX, a.LY)

    def ngx(self):
        return ngx_auto(self.inp.get("NGX", 0), self.inp.get("NGY", 0), self.inp["LX"],
                        self.inp["LY"], self.inp["ENCUT"])

    def energies_hartree(self):
        """Energies fed to the reference-convention solver (m* E, Hartree)."""
        return self.energies_eV * self.mstar / HARTREE


def encut_for(R2, L_ang):
    """ENCUT (eV) with cutoff radius^2 R2 in units of (2 pi/L)^2 (m = m_e, as
    src/grid.f90:401 compares)."""
    L = L_ang / A_0
    return R2 * (2 * math.pi / L) ** 2 / 2 * HARTREE


def qpc_c0(n_energy=400, N_z=200, R2=19.0, grid=64) -> Workload:
    Lx = Ly = 400.0
    dz = 5.0
    LZ = N_z * dz
    ms = MSTAR_GAAS
    x = (np.arange(grid) * Lx / grid)[:, None, None]
    y = (np.arange(grid) * Ly / grid)[None, :, None]
    z = ((np.arange(N_z) + 0.5) * dz)[None, None, :]
    V0, w = 0.010, 80.0
    saddle = V0 * np.exp(-(z - 500.0) ** 2 / (2 * 100.0 ** 2)) * (x - Lx / 2) ** 2 / w ** 2
    hw = 0.002
    # 1/2 m* omega^2 y^2 = (hbar omega)^2 / (2 hbar^2/m*) y^2
    conf = 0.5 * hw ** 2 / (HBAR2_OVER_ME / ms) * (y - Ly / 2) ** 2
    V = (saddle + conf) * np.ones((grid, grid, N_z))
    inp = dict(if_transmission=True, LX=Lx, LY=Ly, LZ=LZ, NKX=1, NKY=1, ENCUT=encut_for(R2, Lx),
               ETA=ms * 1e-6 / HARTREE, TRANSMISSION_GRID=(0.0, ms * 0.020, n_energy))
    E = uniform_energies(0.0, 0.020, n_energy)
    return Workload("qpc_c0", inp, ms * V, E)


def gaussian_disorder(grid_xy, N_z, Lxy, dz, sigma, xi, seed, taper):
    """Gaussian-correlated random field: white noise filtered in Fourier space
    (periodic in x, y; open in z via zero padding), normalised to RMS sigma, tapered
    to zero over `taper` Angstrom at both z ends."""
    rng = np.random.default_rng(seed)
    nzp = N_z + 2 * int(math.ceil(3 * xi / dz))
    noise = rng.standard_normal((grid_xy, grid_xy, nzp))
    kx = 2 * np.pi * np.fft.fftfreq(grid_xy, d=Lxy / grid_xy)
    kz = 2 * np.pi * np.fft.fftfreq(nzp, d=dz)
    K2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    filt = np.exp(-K2 * xi ** 2 / 4)          # sqrt of the Gaussian power spectrum
    f = np.real(np.fft.ifftn(np.fft.fftn(noise) * filt))
    off = (nzp - N_z) // 2
    f = f[:, :, off:off + N_z]
    f *= sigma / np.sqrt(np.mean(f ** 2))
    z = (np.arange(N_z) + 0.5) * dz
    L = N_z * dz
    t = np.clip(np.minimum(z, L - z) / taper, 0.0, 1.0)
    env = np.sin(0.5 * np.pi * t) ** 2
    return f * env[None, None, :]


def disorder_c1(n_energy=4096, N_z=1000, R2=100.5, grid=64, sigma=0.003, seed=20261203) -> Workload:
    Lx = Ly = 400.0
    dz = 2.5
    LZ = N_z * dz
    ms = MSTAR_GAAS
    V = gaussian_disorder(grid, N_z, Lx, dz, sigma, 50.0, seed, 100.0)
    inp = dict(if_transmission=True, LX=Lx, LY=Ly, LZ=LZ, NKX=1, NKY=1, ENCUT=encut_for(R2, Lx),
               ETA=ms * 1e-6 / HARTREE, TRANSMISSION_GRID=(0.0, ms * 0.030, n_energy))
    E = uniform_energies(0.0, 0.030, n_energy)
    return Workload("disorder_c1", inp, ms * V, E)


def qpc_c4_instance(seed, n_energy=512, N_z=400, R2=36.5, grid=32, sigma=(0.0, 0.0005)) -> Workload:
    rng = np.random.default_rng(seed)
    Lx = Ly = 400.0
    LZ = 1000.0
    dz = LZ / N_z
    ms = MSTAR_GAAS
    x = (np.arange(grid) * Lx / grid)[:, None, None]
    y = (np.arange(grid) * Ly / grid)[None, :, None]
    z = ((np.arange(N_z) + 0.5) * dz)[None, None, :]
    V0, w = rng.uniform(0.006, 0.014), rng.uniform(60.0, 100.0)
    z0, zw = rng.uniform(400.0, 600.0), rng.uniform(80.0, 120.0)
    hw = rng.uniform(0.0015, 0.0025)
    saddle = V0 * np.exp(-(z - z0) ** 2 / (2 * zw ** 2)) * (x - Lx / 2) ** 2 / w ** 2
    conf = 0.5 * hw ** 2 / (HBAR2_OVER_ME / ms) * (y - Ly / 2) ** 2
    dis = gaussian_disorder(grid, N_z, Lx, dz, rng.uniform(*sigma), 50.0, seed + 7919, 100.0)
    V = saddle + conf + dis
    inp = dict(if_transmission=True, LX=Lx, LY=Ly, LZ=LZ, NKX=1, NKY=1, ENCUT=encut_for(R2, Lx),
               ETA=ms * 1e-6 / HARTREE, TRANSMISSION_GRID=(0.0, ms * 0.020, n_energy))
    E = uniform_energies(0.0, 0.020, n_energy)
    return Workload("qpc_c4", inp, ms * V, E, extra=dict(V0=V0, w=w, z0=z0, zw=zw, hw=hw))


def disorder_c3(n_energy=1024, N_z=2000, R2=256.5, grid=80, sigma=0.005, seed=20261203) -> Workload:
    """disorder_c1's kind at the largest size: N = 797 (R^2 = 256.5), N_z = 2000 slices of
    0.25 nm, sigma = 5 meV, on an 80 x 80 x N_z grid (the V-hat window of N = 797 needs
    NGX = 35), 1024 energies in 0-30 meV."""
    Lx = Ly = 400.0
    dz = 2.5
    ms = MSTAR_GAAS
    V = gaussian_disorder(grid, N_z, Lx, dz, sigma, 50.0, seed, 100.0)
    inp = dict(if_transmission=True, LX=Lx, LY=Ly, LZ=N_z * dz, NKX=1, NKY=1, ENCUT=encut_for(R2, Lx),
               ETA=ms * 1e-6 / HARTREE, TRANSMISSION_GRID=(0.0, ms * 0.030, n_energy))
    return Workload("disorder_c3", inp, ms * V, uniform_energies(0.0, 0.030, n_energy))


def uneven_c2(seed=20261203, n_devices=48, n_energy=256, sizes=(61, 113, 197, 317, 509), N_z_range=(200, 1500),
              grid=64, sigma=0.003) -> list:
    """The uneven batch: n_devices devices, N drawn from `sizes`, N_z uniform in N_z_range,
    each with its own seeded disorder of disorder_c1's kind (sigma, 5 nm correlation,
    tapered over 10 nm) on a grid x grid x N_z mesh with dz = 0.25 nm, n_energy energies
    in 0-30 meV."""
    R2 = {61: 18.5, 113: 36.5, 197: 64.5, 317: 100.5, 509: 162.5}
    rng = np.random.default_rng(seed)
    out = []
    for d in range(n_devices):
        N = int(rng.choice(sizes))
        N_z = int(rng.integers(N_z_range[0], N_z_range[1] + 1))
        Lx = Ly = 400.0
        dz = 2.5
        ms = MSTAR_GAAS
        V = gaussian_disorder(grid, N_z, Lx, dz, sigma, 50.0, seed + 1000 + d, 100.0)
        inp = dict(if_transmission=True, LX=Lx, LY=Ly, LZ=N_z * dz, NKX=1, NKY=1, ENCUT=encut_for(R2[N], Lx),
                   ETA=ms * 1e-6 / HARTREE, TRANSMISSION_GRID=(0.0, ms * 0.030, n_energy))
        out.append(Workload(f"uneven_c2_{d}", inp, ms * V, uniform_energies(0.0, 0.030, n_energy),
                            extra=dict(N=N, N_z=N_z)))
    return out


def fourier_features(V_eV, shape=(16, 16, 8), scale=0.010 * MSTAR_GAAS):
    """The shape[0] x shape[1] x shape[2] lowest Fourier components of V (eV), as the
    band-limited real samples on a grid of that shape, divided by `scale`."""
    if any(n < s for s, n in zip(shape, V_eV.shape)):
        raise ValueError(f"potential grid {V_eV.shape} is smaller than the feature grid {shape}")
    F = np.fft.fftn(V_eV)
    idx = [np.r_[0:s // 2, n - s // 2:n] for s, n in zip(shape, V_eV.shape)]
    sub = F[np.ix_(*idx)]
    out = np.real(np.fft.ifftn(sub)) * (np.prod(shape) / np.prod(V_eV.shape))
    return (out / scale).astype(np.float32).ravel()

