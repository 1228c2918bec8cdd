"""The FP64 transmission autotuner (NEGF_AUTOTUNE) over the configurations' block sizes:
for each padded size and job count, one context with the flag (every variant timed on
the call's own jobs, the fastest kept) and one without (the fixed rule), on the same
energies.  Prints one JSON line per case: the device seconds of each variant, the
variant tuned and the rule's, and the largest |T| difference between the two.
`python -m negftrans_amd.autotune`."""
from __future__ import annotations

import json

import numpy as np

from . import negf, workloads as W

# (R^2, N_z, job counts): N = 61, 113, 197, 317, 509 and 797
CASES = [(18.5, 200, (64, 256, 400, 800, 1600)), (36.5, 400, (64, 256, 512, 1024)), (64.5, 200, (32, 128, 256, 512)),
         (100.5, 200, (32, 128, 256, 512)), (162.5, 100, (32, 128, 256)), (256.5, 40, (16, 64, 128, 256))]


def disorder_device(R2, N_z, flags):
    """A seeded disorder device of the given basis size (Lx = Ly = 40 nm, 0.25 nm slices,
    sigma = 3 meV, 5 nm correlation), energies 0-30 meV."""
    Lx, dz, ms = 400.0, 2.5, W.MSTAR_GAAS
    grid = 80 if R2 > 200 else 64                  # the V-hat window of N = 797 needs NGX = 35
    V = W.gaussian_disorder(grid, N_z, Lx, dz, 0.003, 50.0, 20261203, 100.0)
    inp = dict(if_transmission=True, LX=Lx, LY=Lx, LZ=N_z * dz, NKX=1, NKY=1, ENCUT=W.encut_for(R2, Lx),
               ETA=ms * 1e-6 / W.HARTREE, TRANSMISSION_GRID=(0.0, ms * 0.030, 4096))
    wl = W.Workload("autotune", inp, ms * V, W.uniform_energies(0.0, 0.030, 4096))
    NGX, NGY = wl.ngx()
    vrec = negf.local_potential_real_to_planewave(wl.V_real_eV / negf.HARTREE, NGX, NGY)
    return negf.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY, flags=flags), wl


def main():
    for R2, N_z, njs in CASES:
        for nj in njs:
            dev, wl = disorder_device(R2, N_z, negf.AUTOTUNE)
            E = wl.energies_hartree()[(np.arange(nj) * (4096 // nj) + 3) % 4096]
            T = dev.transmission(E)
            row = dict(N=dev.N, N_z=N_z, jobs=nj, seconds=dev.tune_seconds(), tuned=dev.last_variant)
            dev.close()
            ref, _ = disorder_device(R2, N_z, 0)
            T2 = ref.transmission(E)
            row.update(rule=ref.last_variant, rule_seconds=ref.last_kernel_seconds,
                       max_dT_vs_rule=float(np.max(np.abs(T - T2))))
            ref.close()
            print(json.dumps(row), flush=True)


if __name__ == "__main__":
    main()
