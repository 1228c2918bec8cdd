"""Throughput of an uneven batch (extension): the devices of workloads.uneven_c2 -- N
from {61, 113, 197, 317, 509}, N_z from 200 to 1500 -- computed

- in one negf_multi hipGraph (all devices at once, longest N_z first per padded size,
  padded sizes on concurrent streams), and
- one device after another with negf_transmission_batch,

on one GPU.  Reports energy-points/s of both (device time from HIP events), the
inversion variant the tuner chose per padded size, and the largest difference between
the two results.  `python -m negftrans_amd.uneven --help`.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import negf, workloads


def measure(n_devices=48, n_energy=32):
    """The comparison of the module docstring, as a dict (progress lines on stdout)."""
    wls = workloads.uneven_c2(n_devices=n_devices, n_energy=n_energy)
    devs, Es = [], []
    for wl in wls:
        NGX, NGY = wl.ngx()
        vrec = negf.local_potential_real_to_planewave(wl.V_real_eV / negf.HARTREE, NGX, NGY)
        devs.append(negf.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY))
        Es.append(wl.energies_hartree())
    print(f"# {len(devs)} devices built: N {sorted(set(d.N for d in devs))}, "
          f"N_z {min(d.N_z for d in devs)}-{max(d.N_z for d in devs)}", flush=True)
    m = negf.Multi(devs)
    t0 = time.time()
    T_multi = m.transmission(Es)
    wall_multi = time.time() - t0
    dev_multi = m.last_seconds
    print(f"# uneven batch: device time {dev_multi:.3f} s, wall {wall_multi:.3f} s", flush=True)
    T_seq, dev_seq = [], 0.0
    t0 = time.time()
    for i, (d, E) in enumerate(zip(devs, Es)):
        T_seq.append(d.transmission(E))
        dev_seq += d.last_kernel_seconds
    wall_seq = time.time() - t0
    print(f"# one device at a time: device time {dev_seq:.3f} s, wall {wall_seq:.3f} s", flush=True)
    pts = sum(len(E) for E in Es)
    diff = max(float(np.max(np.abs(x - y))) for x, y in zip(T_multi, T_seq))
    sizes = sorted(set(-(-d.N // 32) * 32 if d.N <= 512 else d.N for d in devs))
    res = dict(devices=len(devs), energies_per_device=n_energy, energy_points=pts,
               N_values=sorted(set(int(d.N) for d in devs)),
               N_z_range=[int(min(d.N_z for d in devs)), int(max(d.N_z for d in devs))],
               multi=dict(device_s=dev_multi, wall_s=wall_multi, points_per_s=pts / dev_multi),
               sequential=dict(device_s=dev_seq, wall_s=wall_seq, points_per_s=pts / dev_seq),
               speedup=dev_seq / dev_multi, max_abs_dT=diff,
               variants={str(n): negf.Multi.variant(n) for n in sizes if n <= 512})
    m.close()
    for d in devs:
        d.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description="uneven batch throughput (negf_multi vs one device at a time)")
    ap.add_argument("--devices", type=int, default=48)
    ap.add_argument("--energies", type=int, default=32)
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    res = measure(a.devices, a.energies)
    line = json.dumps(res)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
