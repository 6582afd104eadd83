"""Trains the learned troubled-face indicator on the device and writes its weights.

Training data: the faces of FP32 WENO5-Z runs of several flows (3-D Taylor-Green at two Mach numbers and two
meshes, at Mach 0.1 and 64^3 through its transition to turbulence, smooth random Fourier fields in 2-D and 3-D, a
2-D shock interaction, the Kelvin-Helmholtz shear layer through its roll-up, the isentropic vortex, a Sod tube),
sampled at several times. Each
face carries its 15 features and its label (WENO5-Z weight departure above `label`, the same test the audit makes).
Smooth faces, the large majority, are subsampled; faces whose label value lies within a factor 2 of the threshold
are left out of training (they are neither clearly smooth nor clearly troubled); troubled faces are weighted
`pos_weight` in the loss. After training, the decision threshold is set on held-out faces so that at least `recall`
of the troubled ones are caught by the device's own inference (bfloat16 matrix cores); the audit's fallback limit
is 1 % missed.

    python -m fenice_amd.train_indicator [--out fenice_amd/indicator_weights.npz]
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import fvg
from .workloads import kelvin_helmholtz, kh_keys, riemann_states, taylor_green, taylor_green_keys

TWO_PI = "6.283185307179586"

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "indicator_weights.npz")


def _vortex(n, L=10.0, gamma=1.4):
    h = L / n
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    dx = (i.ravel() + 0.5) * h - 0.5 * L
    dy = (j.ravel() + 0.5) * h - 0.5 * L
    r2 = dx * dx + dy * dy
    g = 5.0 / (2.0 * np.pi) * np.exp(0.5 * (1.0 - r2))
    T = 1.0 - (gamma - 1.0) * 25.0 / (8.0 * gamma * np.pi * np.pi) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    u, v, p = 1.0 - g * dy, 1.0 + g * dx, rho * T
    U = np.zeros((n * n, 4))
    U[:, 0], U[:, 1], U[:, 2] = rho, rho * u, rho * v
    U[:, 3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return U


def _shock2d(n, gamma=1.4):
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = (i.ravel() + 0.5) / n
    y = (j.ravel() + 0.5) / n
    q = (x > 0.5).astype(int) + 2 * (y > 0.5).astype(int)   # the four quadrants of a 2-D Riemann problem
    rho = np.array([0.138, 0.5323, 0.5323, 1.5])[q]
    u = np.array([1.206, 0.0, 1.206, 0.0])[q]
    v = np.array([1.206, 1.206, 0.0, 0.0])[q]
    p = np.array([0.029, 0.3, 0.3, 1.5])[q]
    U = np.zeros((n * n, 4))
    U[:, 0], U[:, 1], U[:, 2] = rho, rho * u, rho * v
    U[:, 3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return U


def _modes(n, nd, seed, gamma=1.4, mach=0.3):
    """A smooth random field: a few low Fourier modes in every primitive variable, periodic, n^nd cells."""
    rng = np.random.default_rng(seed)
    axes = np.meshgrid(*[(np.arange(n) + 0.5) / n * 2 * np.pi] * nd, indexing="ij")[::-1]   # x fastest

    def field(scale):
        f = np.zeros(axes[0].shape)
        for _ in range(4):
            k = rng.integers(1, 4, nd)
            f += scale * rng.normal() * np.sin(sum(ki * a for ki, a in zip(k, axes)) + rng.uniform(0, 2 * np.pi))
        return f.ravel()
    rho = 1.0 + np.abs(field(0.1))
    u, v = mach * field(1.0), mach * field(1.0)
    w = mach * field(1.0) if nd == 3 else 0.0 * rho
    p = 1.0 + np.abs(field(0.1))
    U = np.zeros((rho.size, 5 if nd == 3 else 4))
    U[:, 0], U[:, 1], U[:, 2] = rho, rho * u, rho * v
    U[:, 3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    if nd == 3:
        U[:, 4] = rho * w
    return U


def flows():
    """(name, run keys, initial state, snapshot steps)."""
    per2 = dict(bc_xmin="periodic", bc_xmax="periodic", bc_ymin="periodic", bc_ymax="periodic")
    base = dict(gamma="1.4", use_viscous="false", riemann_solver="HLLC", reconstruction="weno5",
                time_scheme="SSP Runge-Kutta 3", cfl="0.5", flux_precision="fp32", nz="1")
    out = []
    for mach in (0.1, 0.6):
        k = dict(taylor_green_keys(64, "fp32"))
        # Mach 0.1 through the transition to turbulence (t = 2.7, 4.5, 7.2 at steps 600, 1000, 1600)
        out.append((f"tgv64-mach{mach}", k, taylor_green(64, mach=mach), (0, 30, 90, 200) + ((600, 1000, 1600) if mach == 0.1 else ())))
    out.append(("tgv128-mach0.1", dict(taylor_green_keys(128, "fp32")), taylor_green(128, mach=0.1), (0, 40)))
    per3 = dict(per2, bc_zmin="periodic", bc_zmax="periodic")
    for seed, (n, nd) in enumerate(((96, 2), (48, 3), (256, 2))):
        keys = dict(base, nx=str(n), ny=str(n), nz=str(n if nd == 3 else 1), Lx=TWO_PI, Ly=TWO_PI, Lz=TWO_PI,
                    **(per3 if nd == 3 else per2))
        out.append((f"modes{n}-{nd}d", keys, _modes(n, nd, seed), (0, 30)))
    out.append(("kh256", dict(kh_keys(256), flux_precision="fp32"), kelvin_helmholtz(256), (0, 150, 400, 700, 1000)))
    out.append(("vortex128", dict(base, nx="128", ny="128", Lx="10.0", Ly="10.0", **per2), _vortex(128), (0, 100)))
    out.append(("riemann2d-192", dict(base, nx="192", ny="192", Lx="1.0", Ly="1.0", bc_xmin="transmissive",
                                      bc_xmax="transmissive", bc_ymin="transmissive", bc_ymax="transmissive"),
                _shock2d(192), (0, 60, 160)))
    sod = riemann_states([400] * 4, np.array([[1.0, 0.125, 1.0, 0.1, 0.0, 0.0, 0.5]] * 4))   # four identical rows
    out.append(("sod400x4", dict(base, nx="400", ny="4", Lx="1.0", Ly="0.01", bc_xmin="transmissive", bc_xmax="transmissive",
                                 bc_ymin="periodic", bc_ymax="periodic"), sod, (0, 100, 250)))
    return out


def collect(label: float, smooth_keep: float, seed: int, log=print, margin: float = 2.0):
    """Features and labels of the flows' faces. Training drops the ambiguous faces whose label value lies within a
    factor `margin` of the label threshold (the audit still judges them by the threshold itself)."""
    rng = np.random.default_rng(seed)
    Xs, ys, tags = [], [], []
    for name, keys, U, snaps in flows():
        cfg = fvg.config_from_keys(keys)
        nd = 3 if int(keys["nz"]) > 1 else 2
        with fvg.Solver(cfg) as s:
            s.upload(U)
            t, done = 0.0, 0
            for k in snaps:
                if k > done:
                    t, n, _ = s.advance(t, 1e9, max_steps=k - done)
                    done = k
                for axis in range(nd):
                    X, d = s.samples(axis)
                    ok = d >= 0
                    X, d = X[ok], d[ok]
                    y = d > label
                    clear = (d > label * margin) | (d < label / margin)
                    keep = clear & (y | (rng.random(d.size) < smooth_keep))
                    Xs.append(X[keep, :15])
                    ys.append(y[keep].astype(np.float32))
                    tags.append(np.full(int(keep.sum()), len(tags)))
                    log(f"  {name} step {k} axis {axis}: {d.size} faces, troubled {int(y.sum())} ({y.mean():.4f})")
    return np.concatenate(Xs), np.concatenate(ys), np.concatenate(tags)


def train(out=DEFAULT_OUT, label=0.05, recall=0.99, epochs=40, seed=1, smooth_keep=0.15, log=print):
    t0 = time.time()
    X, y, _ = collect(label, smooth_keep, seed, log)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(X.shape[0])
    X, y = X[perm], y[perm]
    nval = X.shape[0] // 10
    Xv, yv, Xt, yt = X[:nval], y[:nval], X[nval:], y[nval:]
    pos = max(float(yt.mean()), 1e-6)
    pos_weight = float(min(20.0, 0.25 / pos))
    log(f"samples: train {Xt.shape[0]} (troubled {yt.mean():.4f}), validation {nval}; pos_weight {pos_weight:.2f}")
    p0 = (np.random.default_rng(seed).normal(0.0, 1.0, fvg.MLP_NPARAM) *
          np.concatenate([np.full(480, np.sqrt(2.0 / 15)), np.zeros(32), np.full(1024, np.sqrt(2.0 / 32)), np.zeros(32),
                          np.full(32, np.sqrt(1.0 / 32)), [0.0]])).astype(np.float32)
    with fvg.Trainer(Xt.shape[0]) as tr:
        tr.set_data(Xt, yt)
        l0, _ = tr.grad(p0, pos_weight)
        p, loss = tr.fit(p0, epochs=epochs, batch=1024, lr=2e-3, pos_weight=pos_weight, seed=seed)
        p, loss = tr.fit(p, epochs=max(1, epochs // 3), batch=4096, lr=3e-4, pos_weight=pos_weight, seed=seed + 1)
    z = fvg.mlp_logits(p, Xv)
    zt = np.sort(z[yv > 0.5])
    thr = float(zt[int(np.floor((1.0 - recall) * zt.size))] - 1e-3) if zt.size else 0.0
    pred = z > thr
    rec = float(pred[yv > 0.5].mean()) if zt.size else 1.0
    skip = float((~pred).mean())
    skip_smooth = float((~pred[yv < 0.5]).mean())
    summary = dict(label=label, threshold=thr, loss_start=l0, loss_end=loss, validation_recall=rec, validation_skip=skip,
                   validation_skip_of_smooth=skip_smooth, train_samples=int(Xt.shape[0]), validation_samples=int(nval),
                   troubled_fraction=float(y.mean()), pos_weight=pos_weight, epochs=epochs, seconds=time.time() - t0)
    log(json.dumps(summary))
    np.savez(out, params=p, threshold=np.float32(thr), label=np.float64(label), summary=json.dumps(summary))
    return p, thr, summary


def load(path=DEFAULT_OUT):
    """(params float32[1601], threshold, label, summary dict) of a trained network."""
    d = np.load(path, allow_pickle=False)
    return d["params"], float(d["threshold"]), float(d["label"]), json.loads(str(d["summary"]))


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--label", type=float, default=0.05)
    ap.add_argument("--recall", type=float, default=0.99)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args()
    train(a.out, a.label, a.recall, a.epochs, a.seed, log=lambda m: print(m, flush=True))
