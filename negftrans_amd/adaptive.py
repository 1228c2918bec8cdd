"""Learned adaptive energy grid (extension; NEGFtrans walks a uniform grid).

A TransmissionMLP, trained on T(E) curves that this library's RGF computes, predicts
T(E) on a fixed P-bin grid plus a log-variance per bin from the low-order Fourier
samples of a device potential.  The prediction places the RGF energies where the
predicted curve bends (subband steps, resonances) or is uncertain, and sets how many are
needed.  The RGF is both the fallback and the judge:

- placement: the first RGF mesh has a point density ~ |T''|^(1/3) of the predicted curve
  (the density that minimises the trapezoid error) plus a share that follows the
  predicted uncertainty; adaptive trapezoid refinement on the RGF values then bisects
  the intervals with the largest estimated error until the estimate meets the tolerance;
- fallback: the prediction is compared with the RGF on the first mesh; if they differ by
  more than `fallback_tol` anywhere, the device is run on a uniform grid of n_dense / 8
  points instead;
- judge: the integrated conductance  G = (1 / (E_b - E_a)) int_a^b T(E) dE  (the mean
  transmission over the window, units of 2e^2/h) from the dense n_dense-point RGF grid
  (trapezoid rule) against the same integral over the adaptive points.

`python -m negftrans_amd.adaptive --help` runs the whole experiment (data, training,
judging) on one GPU and prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import negf, workloads
from .mlp import TransmissionMLP


def device_for(wl):
    NGX, NGY = wl.ngx()
    vrec = negf.local_potential_real_to_planewave(wl.V_real_eV / negf.HARTREE, NGX, NGY)
    nx, ny = wl.basis()
    return negf.NegfDevice(nx, ny, wl.atomic(), vrec, NGX, NGY)


def transmission_at(dev, wl, E_eV):
    """RGF T at physical energies E_eV (the m* scaling of the workload applied)."""
    return dev.transmission(np.asarray(E_eV) * wl.mstar / negf.HARTREE)


def integrated(E, T):
    """Mean of T over [E[0], E[-1]] by the trapezoid rule on sorted points."""
    E = np.asarray(E, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    return float(np.sum(0.5 * (T[1:] + T[:-1]) * np.diff(E)) / (E[-1] - E[0]))


def _smooth(x, width):
    """Gaussian smoothing over `width` bins (reflecting ends)."""
    if width <= 0:
        return x
    k = np.arange(-3 * width, 3 * width + 1)
    g = np.exp(-0.5 * (k / width) ** 2)
    g /= g.sum()
    xp = np.concatenate([x[3 * width:0:-1], x, x[-2:-3 * width - 2:-1]])
    return np.convolve(xp, g, mode="valid")


def place(mu, logvar, E_bins, E_lo, E_hi, n0=48, floor=0.2, smooth=3):
    """First RGF mesh of n0 points on [E_lo, E_hi] from the predicted curve: the quantiles
    of a density ~ |T''|^(1/3) (the density that minimises the trapezoid error for a given
    count), taken from the smoothed predicted mean, plus a share following the predicted
    standard deviation and a uniform floor.  Both window ends are included."""
    mu = _smooth(np.asarray(mu, dtype=np.float64), smooth)
    E_bins = np.asarray(E_bins, dtype=np.float64)
    P = len(mu)
    d2 = np.zeros(P)
    d2[1:-1] = np.abs(mu[2:] - 2 * mu[1:-1] + mu[:-2])
    curv = np.cbrt(d2)
    sig = np.exp(0.5 * np.clip(np.asarray(logvar, dtype=np.float64), -20.0, 20.0))
    dens = curv / max(curv.sum(), 1e-300) + 0.5 * sig / max(sig.sum(), 1e-300) + floor / P
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]))])
    cdf /= cdf[-1]
    E = np.interp(np.linspace(0.0, 1.0, n0), cdf, E_bins)
    E[0], E[-1] = E_lo, E_hi
    return np.unique(np.clip(E, E_lo, E_hi))


def refine(dev, wl, E, T, tol, n_max, rounds=10):
    """Adaptive trapezoid quadrature on RGF values: per interval the error is estimated as
    h^3 |T''| / 12 (T'' from the three-point formula at either end, the larger), and the
    intervals with the largest estimates are bisected -- all new midpoints of a round in
    one batched RGF call -- until the estimated error of the mean over the window is
    below tol / 2 or n_max points are used.  Returns (E, T)."""
    E = np.asarray(E, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    width = E[-1] - E[0]
    for _ in range(rounds):
        h = np.diff(E)
        c = np.zeros(len(h))
        if len(E) >= 3:
            f2 = np.abs(2 * ((T[2:] - T[1:-1]) / h[1:] - (T[1:-1] - T[:-2]) / h[:-1]) / (h[1:] + h[:-1]))
            c[:-1] = np.maximum(c[:-1], f2)
            c[1:] = np.maximum(c[1:], f2)
        err = h ** 3 * c / 12.0 / width
        if err.sum() <= 0.5 * tol or len(E) >= n_max:
            break
        order = np.argsort(-err)
        keep = np.cumsum(err[order])
        # bisecting an interval cuts its estimate by about 4: pick the largest ones until
        # what is left (plus a quarter of the picked) is under tol / 4
        left = err.sum() - keep + 0.25 * keep
        k = int(np.searchsorted(-left, -0.25 * tol)) + 1
        k = max(1, min(k, len(order), n_max - len(E)))
        pick = np.sort(order[:k])
        mids = 0.5 * (E[pick] + E[pick + 1])
        Tm = transmission_at(dev, wl, mids)
        E = np.concatenate([E, mids])
        T = np.concatenate([T, Tm])
        o = np.argsort(E)
        E, T = E[o], T[o]
    return E, T


def adaptive_conductance(dev, wl, mu, logvar, E_bins, E_lo, E_hi, n_dense, tol=1e-4, n0=48,
                         fallback_tol=0.25):
    """(G, RGF energies used, fell back?) for one device: first mesh from the prediction,
    then refine().  Where the RGF on the first mesh differs from the prediction by more
    than fallback_tol the prediction is not trusted and the device is run on a uniform
    grid of n_dense / 8 points instead."""
    E = place(mu, logvar, E_bins, E_lo, E_hi, n0)
    T = transmission_at(dev, wl, E)
    if np.max(np.abs(T - np.interp(E, E_bins, mu))) > fallback_tol:
        Ef = np.linspace(E_lo, E_hi, n_dense // 8)
        return integrated(Ef, transmission_at(dev, wl, Ef)), len(Ef), True
    E, T = refine(dev, wl, E, T, tol, n_dense // 8)
    return integrated(E, T), len(E), False


def dataset(seeds, n_bins, log_every=0, **kw):
    X, T = [], []
    for i, s in enumerate(seeds):
        if log_every and i % log_every == 0:
            print(f"# data: device {i}", flush=True)
        wl = workloads.qpc_c4_instance(int(s), n_energy=n_bins, **kw)
        X.append(workloads.fourier_features(wl.V_real_eV))
        dev = device_for(wl)
        T.append(transmission_at(dev, wl, wl.energies_eV))
        dev.close()
    return np.array(X, dtype=np.float32), np.array(T, dtype=np.float32)


def uniform_points_needed(dev, wl, G_ref, tol, E_lo, E_hi, n_start=65, n_max=8193):
    """The smallest uniform grid of n = 64 * 2^j + 1 points on [E_lo, E_hi] (j = 0, 1, ...,
    up to n_max) whose trapezoid G is within tol of G_ref; n_max + 1 if none is."""
    n = n_start
    while n <= n_max:
        Eu = np.linspace(E_lo, E_hi, n)
        if abs(integrated(Eu, transmission_at(dev, wl, Eu)) - G_ref) <= tol:
            return n
        n = 2 * n - 1
    return n_max + 1


def judge(model, seeds, n_bins, n_dense, tol=1e-4, n0=48, fallback_tol=0.25, log_every=0, uniform_search=False, **kw):
    """Per held-out device: the dense G; the network-seeded adaptive G and its RGF count;
    the same refinement started from a uniform mesh of n0 points (the method without the
    network) and its count; a uniform grid with the network-seeded count; with
    uniform_search, also the smallest uniform grid (64 * 2^j + 1 points) that meets tol."""
    rows = []
    for i, s in enumerate(seeds):
        if log_every and i % log_every == 0:
            print(f"# judge: device {i}", flush=True)
        wl = workloads.qpc_c4_instance(int(s), n_energy=n_bins, **kw)
        dev = device_for(wl)
        Ed = workloads.uniform_energies(0.0, 0.020, n_dense)
        G_dense = integrated(Ed, transmission_at(dev, wl, Ed))
        E_lo, E_hi = Ed[0], Ed[-1]
        mu, lv = model.predict(workloads.fourier_features(wl.V_real_eV)[None, :])
        G_ad, n_pts, fb = adaptive_conductance(dev, wl, mu[0], lv[0], wl.energies_eV, E_lo, E_hi, n_dense, tol, n0,
                                               fallback_tol)
        Eu0 = np.linspace(E_lo, E_hi, n0)
        Eb, Tb = refine(dev, wl, Eu0, transmission_at(dev, wl, Eu0), tol, n_dense // 8)
        Eu = np.linspace(E_lo, E_hi, n_pts)
        rows.append(dict(seed=int(s), G_dense=G_dense, G_adaptive=G_ad, points=n_pts, fallback=bool(fb),
                         G_refine_uniform_start=integrated(Eb, Tb), points_uniform_start=len(Eb),
                         G_uniform=integrated(Eu, transmission_at(dev, wl, Eu))))
        if uniform_search:
            rows[-1]["uniform_points_needed"] = uniform_points_needed(dev, wl, G_dense, tol, E_lo, E_hi)
        dev.close()
    return rows


def summary(rows, n_dense):
    def err(key):
        return np.array([abs(r[key] - r["G_dense"]) for r in rows])
    ea, eb, eu = err("G_adaptive"), err("G_refine_uniform_start"), err("G_uniform")
    pa = np.array([r["points"] for r in rows])
    pb = np.array([r["points_uniform_start"] for r in rows])
    out = {}
    if rows and "uniform_points_needed" in rows[0]:
        pu = np.array([r["uniform_points_needed"] for r in rows])
        out = dict(uniform_needed_mean=float(pu.mean()), uniform_needed_max=int(pu.max()))
    return dict(out, devices=len(rows), dense_points=n_dense, fallback_rate=float(np.mean([r["fallback"] for r in rows])),
                points_mean=float(pa.mean()), points_max=int(pa.max()),
                err_median=float(np.median(ea)), err_max=float(ea.max()), within_1e_4=float(np.mean(ea <= 1e-4)),
                no_network_points_mean=float(pb.mean()), no_network_err_median=float(np.median(eb)),
                no_network_err_max=float(eb.max()),
                uniform_same_points_err_median=float(np.median(eu)), uniform_same_points_err_max=float(eu.max()))


def main(argv=None):
    ap = argparse.ArgumentParser(description="learned adaptive energy grid: data, training, judging")
    ap.add_argument("--train", type=int, default=512)
    ap.add_argument("--test", type=int, default=16)
    ap.add_argument("--bins", type=int, default=512)
    ap.add_argument("--dense", type=int, default=8192)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--n0", type=int, default=48)
    ap.add_argument("--N_z", type=int, default=400)
    ap.add_argument("--R2", type=float, default=36.5)
    ap.add_argument("--sigma", type=float, nargs=2, default=(0.0, 0.0005),
                    help="range (eV) of the disorder strength drawn per device")
    ap.add_argument("--uniform-search", action="store_true",
                    help="also find the smallest uniform grid that meets the tolerance")
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    kw = dict(N_z=a.N_z, R2=a.R2, sigma=tuple(a.sigma))
    t0 = time.time()
    X, T = dataset(range(1000, 1000 + a.train), a.bins, log_every=64, **kw)
    t1 = time.time()
    print(f"# training data: {a.train} devices x {a.bins} energies in {t1 - t0:.1f} s", flush=True)
    model = TransmissionMLP(X.shape[1], a.bins, seed=0)
    hist = model.fit(X, T, epochs=a.epochs, batch=64, lr=1e-3, log_every=max(1, a.epochs // 10))
    t2 = time.time()
    print(f"# training: {a.epochs} epochs in {t2 - t1:.1f} s, loss {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    rows = judge(model, range(900000, 900000 + a.test), a.bins, a.dense, a.tol, a.n0, log_every=4,
                 uniform_search=a.uniform_search, **kw)
    t3 = time.time()
    res = dict(summary(rows, a.dense), train_devices=a.train, epochs=a.epochs, loss_first=hist[0],
               loss_last=hist[-1], seconds=dict(data=t1 - t0, train=t2 - t1, judge=t3 - t2), rows=rows)
    line = json.dumps(res)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
