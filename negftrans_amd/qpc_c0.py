"""Transmission of NEGFtrans's own case (N = 61, N_z = 200, the 400 energies of 0-20 meV,
m* scaled) on one GPU, FP64.  Prints one JSON line per repeat: HIP-event device time of
the call and energy-points/s.  A call of 400 jobs is few enough for two workgroups per
job on this block size, so it runs the two-sided form.
`python -m negftrans_amd.qpc_c0 [--repeats 3] [--energies 400]`."""
from __future__ import annotations

import argparse
import json

import numpy as np

from . import negf, workloads


def measure(repeats=3, n_energy=400, log=None):
    """One dict per repeat (see the module docstring); the first call is a warm-up."""
    wl = workloads.qpc_c0()
    NGX, NGY = wl.ngx()
    vrec = negf.local_potential_real_to_planewave(wl.V_real_eV / negf.HARTREE, NGX, NGY)
    dev = negf.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY)
    Eall = wl.energies_hartree()
    E = Eall[np.arange(n_energy) % len(Eall)]
    dev.transmission(E)
    out = []
    for _ in range(repeats):
        T = dev.transmission(E)
        t = dev.last_kernel_seconds
        out.append(dict(N=dev.N, N_z=dev.N_z, energies=len(E), device_s=t, points_per_s=len(E) / t,
                        T_max=float(T.max()), status=dev.last_status))
        if log:
            log(json.dumps(out[-1]))
    dev.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="N = 61, N_z = 200 transmission timing")
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--energies", type=int, default=400)
    a = ap.parse_args(argv)
    measure(a.repeats, a.energies, log=lambda m: print(m, flush=True))


if __name__ == "__main__":
    main()
