"""Equilibrium density of the metric configuration (N = 317, N_z = 1000, disorder_c1)
with 64 contour nodes on the arc (the reference's Simpson rule, N_circle = 64, by
default; Gauss-Legendre as the extension), as one negf_density_batch call on one GPU.
Prints one JSON line (device time from HIP events, nodes/s, status).
`python -m negftrans_amd.density_c1 [--nodes 64] [--quadrature simpson|gauss-legendre]`."""
from __future__ import annotations

import argparse
import json
import time

from . import negf, workloads


def measure(nodes=64, repeats=1, log=print, quadrature="simpson"):
    """Device time of the density call (best of `repeats`), as a dict."""
    wl = workloads.disorder_c1()
    NGX, NGY = wl.ngx()
    vrec = negf.local_potential_real_to_planewave(wl.V_real_eV / negf.HARTREE, NGX, NGY)
    at0 = wl.atomic()
    dev = negf.NegfDevice(*wl.basis(), at0, vrec, NGX, NGY)
    # Fermi level at 15 meV (m* scaled, as the workload's energies), 4 K
    at = negf.Atomic(V_L=at0.V_L, V_R=at0.V_R, ETA=at0.ETA, LX=at0.LX, LY=at0.LY, LZ=at0.LZ, ENCUT=at0.ENCUT,
                     MU=wl.mstar * 0.015 / negf.HARTREE, TEMPERATURE=4.0, GAP=6 * negf.K_B * 4.0)
    z, w = negf.density_contour(at, nodes, 0, float(wl.V_real_eV.min() / negf.HARTREE), quadrature)
    times = []
    for r in range(repeats):
        t0 = time.time()
        dev.density_matrix(z, w)
        times.append((dev.last_kernel_seconds, time.time() - t0))
        log(f"# repeat {r}: device {times[-1][0]:.3f} s")
    best = min(t[0] for t in times)
    res = dict(N=dev.N, N_z=dev.N_z, nodes=len(z), quadrature=quadrature, device_s=best,
               nodes_per_s=len(z) / best, status=dev.last_status)
    dev.close()
    return res


def measure_bias(nodes=32, vds_meV=4.0, log=print):
    """Device time of the bias-window blocks (negf_neq_density_batch) of the same device:
    `nodes` Gauss-Legendre nodes over [mu_R - GAP, MU + GAP] at VDS = vds_meV (m* scaled as
    the workload's energies), MU = 15 meV, 4 K."""
    wl = workloads.disorder_c1()
    NGX, NGY = wl.ngx()
    vrec = negf.local_potential_real_to_planewave(wl.V_real_eV / negf.HARTREE, NGX, NGY)
    dev = negf.NegfDevice(*wl.basis(), wl.atomic(), vrec, NGX, NGY)
    ha = negf.Atomic(MU=wl.mstar * 0.015 / negf.HARTREE, TEMPERATURE=4.0, GAP=6 * negf.K_B * 4.0)
    E, w = negf.bias_window(ha, wl.mstar * vds_meV * 1e-3 / negf.HARTREE, nodes, "gauss-legendre")
    t0 = time.time()
    dev.neq_density_matrix(E, w)
    res = dict(N=dev.N, N_z=dev.N_z, bias_nodes=len(E), device_s=dev.last_kernel_seconds,
               wall_s=time.time() - t0, nodes_per_s=len(E) / dev.last_kernel_seconds, status=dev.last_status)
    log(f"# bias window: {len(E)} nodes, device {res['device_s']:.3f} s")
    dev.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description="64-node contour density at N = 317, N_z = 1000")
    ap.add_argument("--nodes", type=int, default=64)
    ap.add_argument("--repeats", type=int, default=1)
    ap.add_argument("--quadrature", choices=("simpson", "gauss-legendre"), default="simpson")
    ap.add_argument("--bias-nodes", type=int, default=0,
                    help="instead: time the bias-window blocks with this many nodes (VDS = 4 meV)")
    a = ap.parse_args(argv)
    if a.bias_nodes:
        res = measure_bias(a.bias_nodes, log=lambda m: print(m, flush=True))
    else:
        res = measure(a.nodes, a.repeats, log=lambda m: print(m, flush=True), quadrature=a.quadrature)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
