"""Transmission at the largest configuration's block size (N = 797, padded 800: the
four-tile matrix-core Gauss-Jordan) on its disorder, at a reduced depth N_z, for a
batch of energies on one GPU.  Prints one JSON line per batch size: HIP-event device
time, time per slice per job, and the energy-points/s that rate gives at N_z = 2000.
`python -m negftrans_amd.block_c3 [--nz 40] [--energies 64 256]`."""
from __future__ import annotations

import argparse
import json

import numpy as np

from . import negf, workloads


def measure(nz=40, batches=(64, 256), log=None):
    """One dict per batch size (see the module docstring)."""
    wl = workloads.disorder_c3(N_z=nz)
    NGX, NGY = wl.ngx()
    vrec = negf.local_potential_real_to_planewave(wl.V_real_eV / negf.HARTREE, NGX, NGY)
    dev = negf.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY)
    Eall = wl.energies_hartree()
    dev.transmission(Eall[:8])                   # warm-up
    out = []
    for n in batches:
        E = Eall[(np.arange(n) * 3 + 500) % len(Eall)]
        T = dev.transmission(E)
        t = dev.last_kernel_seconds
        out.append(dict(N=dev.N, N_z=dev.N_z, energies=n, device_s=t, ms_per_slice_per_job=t / dev.N_z * 1e3,
                        points_per_s_at_Nz2000=n / (t / dev.N_z * 2000), T_max=float(T.max()),
                        status=dev.last_status))
        if log:
            log(json.dumps(out[-1]))
    dev.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="N = 797 transmission timing")
    ap.add_argument("--nz", type=int, default=40)
    ap.add_argument("--energies", type=int, nargs="+", default=[64, 256])
    a = ap.parse_args(argv)
    measure(a.nz, a.energies, log=lambda m: print(m, flush=True))


if __name__ == "__main__":
    main()
