"""Benchmark of the uniform-grid residual and time step on one MI355X.

Workload: the 3-D Taylor-Green vortex (Navier-Stokes, periodic [0, 2 pi]^3, 256^3 cells, FP64, Mach 0.1, Re 1600,
Pr 0.71, rho0 = 1, u = sin x cos y cos z, v = -cos x sin y cos z, w = 0,
p = p0 + (cos 2x + cos 2y)(cos 2z + 2)/16 with p0 = 1/(gamma Ma^2)), WENO5-Z + HLLC + SSP-RK3, CFL 0.5.
A "step" is one SSP-RK3 step (three residual stages) of the whole mesh; the metric counts cell updates per
stage: value = cells * 3 * steps / seconds / 1e6.

    python bench.py [--steps K] [--warmup W] [--n 256]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, published
FP64_VECTOR_PEAK_TFLOPS = 78.6  # MI355X FP64 vector, published (FMA counted as 2)
BYTES_PER_CELL_STAGE_RK3_3D = 8.0 * 5 * (2 + 3 + 3) / 3.0   # 106.7 B: stage input, U^n on stages 2-3, output


def tgv_state(n: int, gamma: float = 1.4, mach: float = 0.1) -> np.ndarray:
    h = 2.0 * math.pi / n
    x = (np.arange(n) + 0.5) * h
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    u = np.sin(x3) * np.cos(y3) * np.cos(z3)
    v = -np.cos(x3) * np.sin(y3) * np.cos(z3)
    p = 1.0 / (gamma * mach * mach) + (np.cos(2 * x3) + np.cos(2 * y3)) * (np.cos(2 * z3) + 2.0) / 16.0
    U = np.empty((n ** 3, 5))
    U[:, 0] = 1.0
    U[:, 1] = u.ravel()
    U[:, 2] = v.ravel()
    U[:, 3] = p.ravel() / (gamma - 1.0) + 0.5 * (u * u + v * v).ravel()
    U[:, 4] = 0.0
    return U


def run_keys(n: int, flux: str = "fp64") -> dict:
    L = "6.283185307179586"
    return dict(nx=n, ny=n, nz=n, Lx=L, Ly=L, Lz=L, gamma=1.4, use_viscous="true", mu="6.25e-4", pr="0.71",
                bc_xmin="periodic", bc_xmax="periodic", bc_ymin="periodic", bc_ymax="periodic", bc_zmin="periodic",
                bc_zmax="periodic", riemann_solver="HLLC", reconstruction="weno5", time_scheme="SSP Runge-Kutta 3", cfl="0.5",
                flux_precision=flux)


def error_vs_fp64(n: int, flux: str, steps: int = 20) -> dict:
    """L1 / Linf relative error of the density of a `flux` run against the FP64 path (bit-identical to the oracle), both
    from the Taylor-Green state at the same fixed dt (below the CFL step), after `steps` steps."""
    from fenice_amd import fvg
    U0 = tgv_state(n)
    out = {}
    for f in ("fp64", flux):
        with fvg.Solver(fvg.config_from_keys(run_keys(n, f))) as s:
            s.upload(U0)
            t = 0.0
            for _ in range(steps):
                t += s.step(t, 1e9, 1e-3).dt
            out[f] = s.download()[:, 0]
    d = np.abs(out[flux] - out["fp64"])
    return {"l1_rel": float(d.sum() / np.abs(out["fp64"]).sum()), "linf_rel": float(d.max() / np.abs(out["fp64"]).max()),
            "steps": steps, "dt": 1e-3, "mesh": n}


# ---------------------------------------------------------------------------------------------------------------
# CPU baseline (runs in a child process): the reference's own code built from its sources (oracle/_ref), or the
# oracle where that build is absent, on the same workload, a bounded sample of steps.
# ---------------------------------------------------------------------------------------------------------------
def cpu_baseline_worker(n: int, budget_s: float) -> dict:
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libfenice_ref.so")
    U0 = tgv_state(n)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    if os.path.exists(ref_so):
        L = C.CDLL(ref_so)
        L.ref_create.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p)]
        L.ref_create.restype = C.c_void_p
        L.ref_set_state.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.ref_step.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        keys = run_keys(n)
        ks = [k.encode() for k in keys]
        vs = [str(v).encode() for v in keys.values()]
        # a run description of our own for the reference's case loader (every key of the run is an override below)
        import tempfile
        case_dir = tempfile.mkdtemp(prefix="fvg_case_")
        case_file = os.path.join(case_dir, "case.toml")
        with open(case_file, "w") as f:
            f.write('[case]\nname = "Uniform run"\n\n[initial]\nstate = { rho = 1.0, u = 0.0, v = 0.0, p = 1.0 }\n')
        h = L.ref_create(case_file.encode(), len(ks), (C.c_char_p * len(ks))(*ks), (C.c_char_p * len(vs))(*vs))
        if not h:
            L.ref_last_error.restype = C.c_char_p
            raise RuntimeError(L.ref_last_error().decode())
        L.ref_set_state(h, U0.ctypes.data_as(C.POINTER(C.c_double)))
        dt, s = C.c_double(), C.c_double()

        def step(t):
            L.ref_step(h, t, 1e9, 0.0, C.byref(dt), C.byref(s))
            return dt.value
        kind = "reference"
    else:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import Oracle
        from fenice_amd.fvg import config_from_keys
        ora = Oracle()
        cfg = config_from_keys(run_keys(n))
        U = U0.copy()

        def step(t):
            return ora.step(cfg, U, t, 1e9)[0]
        kind = "port"
    t = 0.0
    t0 = time.perf_counter()
    t += step(t)   # first step: first-touch of every work array, not timed
    warm = time.perf_counter() - t0
    steps = max(1, min(3, int(budget_s / max(warm, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(steps):
        t += step(t)
    el = time.perf_counter() - t0
    return {"value": n ** 3 * 3 * steps / el / 1e6, "unit": "Mcell-updates/s per RK stage", "cores": threads, "kind": kind,
            "sample": f"3-D Taylor-Green {n}^3 WENO5-Z+HLLC+SSP-RK3 NS, {steps} timed step(s) after 1 untimed warm-up step "
                      f"({el:.1f} s timed), OpenMP {threads} threads" + (", AVX2 kernels" if kind == "reference" else "")}


def cpu_baseline(n: int, budget_s: float):
    env = dict(os.environ)
    env.pop("FENICE_NO_SIMD", None)
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-worker", "--n", str(n), "--cpu-budget", str(budget_s)],
                       capture_output=True, text=True, env=env)
    if r.returncode != 0:
        return {"value": None, "unit": "Mcell-updates/s per RK stage", "cores": None, "kind": None,
                "sample": "failed: " + (r.stderr or r.stdout)[-300:]}
    return json.loads(r.stdout.strip().splitlines()[-1])


def copy_calibration(nbytes: int = 1 << 30) -> None:
    """One device-to-device copy of known size (HIP's copy kernel), so that a counter run can scale FETCH_SIZE /
    WRITE_SIZE: the copy moves nbytes in and nbytes out."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    hip.hipMalloc(C.byref(a), C.c_size_t(nbytes))
    hip.hipMalloc(C.byref(b), C.c_size_t(nbytes))
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-n", type=int, default=0, help="mesh of the CPU sample (default: the GPU mesh)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-worker", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true", help="end with a 1 GiB device copy (counter calibration)")
    ap.add_argument("--flux", default="fp64", choices=["fp64", "fp32"], help="arithmetic of reconstruction and fluxes")
    a = ap.parse_args()
    if a.cpu_worker:
        print(json.dumps(cpu_baseline_worker(a.n, a.cpu_budget)))
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {a.gpus}); "
              "see DESIGN.md for how the mesh would be split across GPUs", file=sys.stderr)
        sys.exit(2)

    from fenice_amd import fvg
    n = a.n
    cfg = fvg.config_from_keys(run_keys(n, a.flux))
    U0 = tgv_state(n)
    cells = n ** 3
    with fvg.Solver(cfg) as s:
        s.upload(U0)
        t = 0.0
        if a.warmup > 0:
            t, _, _ = s.advance(t, 1e9, max_steps=a.warmup)
        tot0 = s.conserved_totals()
        s.set_timing(True)
        t0 = time.perf_counter()
        t, steps, s_last = s.advance(t, 1e9, max_steps=a.steps)
        el = time.perf_counter() - t0
        times = s.kernel_times()
        tot1 = s.conserved_totals()
    stages = 3 * steps
    value = cells * stages / el / 1e6
    # the stage as the unit: algorithmic bytes of one stage over the event-measured time of its kernels
    fam = {k: v for k, v in times.items() if v[1] > 0}
    stage_ms = sum(ms for ms, _ in fam.values()) / stages
    achieved = BYTES_PER_CELL_STAGE_RK3_3D * cells / (stage_ms * 1e-3) / 1e9
    dom = max(fam.items(), key=lambda kv: kv[1][0])
    out = {
        "metric": "Mcell-updates/s per RK stage, % HBM peak, at stated L1 error vs FP64 oracle",
        "value": round(value, 3), "unit": "Mcell-updates/s per RK stage", "n_gpus": 1, "steps": steps, "warmup": a.warmup,
        "ms_per_step": round(el / steps * 1e3, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.flux == "fp64" else "f32 fluxes, f64 state", "data": "synthetic (analytic Taylor-Green initial field)",
        "config": {"workload": f"3-D Taylor-Green vortex {n}^3, Navier-Stokes Re 1600 Pr 0.71 Mach 0.1, WENO5-Z + HLLC + SSP-RK3, "
                               "CFL 0.5, periodic", "cells": cells, "stages_per_step": 3},
        "l1_error_vs_fp64_oracle": 0.0 if a.flux == "fp64" else None,
        "roofline": {"bound": "hbm", "kernel": "one RK stage (k_cells, k_faces x/y/z, k_control, k_update)",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None, "bytes_per_cell_stage": round(BYTES_PER_CELL_STAGE_RK3_3D, 2)},
        "kernels": {k: {"ms_total": round(ms, 3), "launches": cnt, "avg_ms": round(ms / cnt, 4)} for k, (ms, cnt) in fam.items()},
        "dominant_kernel": dom[0],
        "conservation": {"mass_rel_drift": abs(tot1[0] - tot0[0]) / abs(tot0[0]), "energy_rel_drift": abs(tot1[2] - tot0[2]) / abs(tot0[2])},
    }
    if a.flux != "fp64":
        out["error_vs_fp64"] = error_vs_fp64(min(n, 128), a.flux)
        out["l1_error_vs_fp64_oracle"] = out["error_vs_fp64"]["l1_rel"]
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.cpu_n or n, a.cpu_budget)
    print(json.dumps(out), flush=True)
    if a.calibrate_copy:
        copy_calibration()


if __name__ == "__main__":
    main()
