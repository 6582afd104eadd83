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
# HBM bytes of one stage of the 256^3 workload (k_cells + k_faces x/y/z + k_update, the update averaged over the
# three SSP-RK3 stages), from two rocprofv3 --pmc runs (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0, the factors a 1 GiB
# device copy reads in the same runs); the runs and per-kernel figures are in the profile files named here. These
# are figures of those counter runs, not of the run that prints them.
MEASURED_STAGE_TRAFFIC = {("fp64", 256): (16.888e9, "profiles/tgv256_weno5_rk3_fp64.md"),
                          ("fp32", 256): (7.678e9, "profiles/tgv256_weno5_rk3_fp32.md")}
# VALU instructions per wavefront (SQ_INSTS_VALU / SQ_WAVES, one face or cell per lane) of each kernel of the
# workload, from the same counter runs; with this run's event time of the kernel they give its VALU issue rate. A
# wave64 VALU instruction occupies a SIMD for 4 cycles; MI355X has 1,024 SIMDs at up to 2.4 GHz.
VALU_PER_WAVE = {"fp64": ({"k_cells": 578, "k_faces_x": 1893, "k_faces_y": 1902, "k_faces_z": 1873}, "profiles/tgv256_weno5_rk3_fp64.md")}
SIMDS, CLOCK_HZ = 1024, 2.4e9

from fenice_amd.workloads import taylor_green as tgv_state  # noqa: E402
from fenice_amd.workloads import taylor_green_keys as run_keys  # noqa: E402


PASSES = "default"   # kernel sequence of the residual: the context's default, "fused" or "separate" (--passes)


def make_solver(n: int, flux: str, indicator: str = "none", weights: str = ""):
    """The solver of the workload; with indicator = learned, the trained network's weights are loaded."""
    from fenice_amd import fvg
    keys = dict(run_keys(n, flux))
    if indicator == "learned":
        keys.update(indicator="learned", indicator_audit="50")
    s = fvg.Solver(fvg.config_from_keys(keys))
    if PASSES != "default":
        s.set_fused(PASSES == "fused")
    if indicator == "learned":
        from fenice_amd.train_indicator import DEFAULT_OUT, load
        params, thr, _, _ = load(weights or DEFAULT_OUT)
        s.set_indicator(params, thr)
    return s


def error_vs_fp64(n: int, flux: str, indicator: str = "none", weights: str = "", steps: int = 20) -> dict:
    """L1 / Linf relative error of the density of a `flux` (and `indicator`) run against the FP64 path (bit-identical
    to the oracle), both from the Taylor-Green state at the same fixed dt (below the CFL step), after `steps` steps."""
    U0 = tgv_state(n)
    out = []
    for f, ind in (("fp64", "none"), (flux, indicator)):
        with make_solver(n, f, ind, weights) as s:
            s.upload(U0)
            t = 0.0
            for _ in range(steps):
                t += s.step(t, 1e9, 1e-3).dt
            out.append(s.download()[:, 0])
    d = np.abs(out[1] - out[0])
    return {"l1_rel": float(d.sum() / np.abs(out[0]).sum()), "linf_rel": float(d.max() / np.abs(out[0]).max()),
            "steps": steps, "dt": 1e-3, "mesh": n}


# ---------------------------------------------------------------------------------------------------------------
# CPU baseline (runs in a child process): the reference's own code built from its sources (oracle/_ref), or the
# oracle where that build is absent, on the same workload, a bounded sample of steps.
# ---------------------------------------------------------------------------------------------------------------
def cpu_workload(name: str, n: int, quadrature: str = "midpoint"):
    """(keys, initial state (cells, 5), cells, RK stages per step, most timed steps, description) of a CPU sample."""
    from fenice_amd.workloads import kelvin_helmholtz, kh_keys, vortex_keys, vortex_state

    def five(U4):
        U = np.zeros((U4.shape[0], 5))
        U[:, :4] = U4
        return U
    if name == "vortex":
        return (vortex_keys(n, cfl="0.7"), five(vortex_state(n)), n * n, 1, 200,
                f"isentropic vortex {n}^2 MUSCL Van Leer+HLLC+MUSCL-Hancock, CFL 0.7")
    if name == "kh":
        return kh_keys(n), five(kelvin_helmholtz(n)), n * n, 3, 3, f"Kelvin-Helmholtz {n}^2 WENO5-Z+HLLC+SSP-RK3"
    return (run_keys(n, "fp64", quadrature), tgv_state(n), n ** 3, 3, 12,
            f"3-D Taylor-Green {n}^3 WENO5-Z+HLLC+SSP-RK3 NS" + (", fourth-order flux quadrature" if quadrature == "fourth" else ""))


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or '' where it is not given."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def cpu_baseline_worker(n: int, budget_s: float, quadrature: str = "midpoint", workload: str = "tgv", dump: str = "") -> dict:
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libfenice_ref.so")
    keys, U0, cells, nst, cap, desc = cpu_workload(workload, n, quadrature)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    if os.path.exists(ref_so):
        L = C.CDLL(ref_so)
        L.ref_create.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p)]
        L.ref_create.restype = C.c_void_p
        L.ref_set_state.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.ref_step.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
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
        L.ref_get_state.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        dt, s = C.c_double(), C.c_double()

        def step(t):
            L.ref_step(h, t, 1e9, 0.0, C.byref(dt), C.byref(s))
            return dt.value

        def state():
            out = np.empty_like(U0)
            L.ref_get_state(h, out.ctypes.data_as(C.POINTER(C.c_double)))
            return out
        kind = "reference"
    else:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import Oracle
        from fenice_amd.fvg import config_from_keys
        ora = Oracle()
        cfg = config_from_keys(keys)
        U = U0.copy()
        lag = [0.0]

        def step(t):
            dt, _, lag[0], _ = ora.step(cfg, U, t, 1e9, 0.0, lag[0])
            return dt

        def state():
            return U
        kind = "port"
    t = 0.0
    t0 = time.perf_counter()
    t += step(t)   # first step: first-touch of every work array, not timed
    warm = time.perf_counter() - t0
    steps = max(1, min(cap, int(budget_s / max(warm, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(steps):
        t += step(t)
    el = time.perf_counter() - t0
    out = {"value": cells * nst * steps / el / 1e6, "unit": "Mcell-updates/s per RK stage", "cores": threads, "kind": kind,
           "cpu_model": cpu_model(),
           "sample": desc + f", {steps} timed step(s) after 1 untimed warm-up step "
                     f"({el:.1f} s timed), OpenMP {threads} threads" + (", AVX2 kernels" if kind == "reference" else "")}
    if dump:   # the state after every step taken (warm-up and timed), for the device's comparison
        np.save(dump, state())
        out["state_steps"], out["state_t"] = steps + 1, t
    return out


_ENS = {}


def _ens_init():
    """A worker of the ensemble sample: the reference's library and the ensemble's instances."""
    from fenice_amd.workloads import ensemble_keys, riemann_ensemble
    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libfenice_ref.so"))
    L.ref_create.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p)]
    L.ref_create.restype = C.c_void_p
    L.ref_set_state.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.ref_step.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.ref_destroy.argtypes = [C.c_void_p]
    import tempfile
    case_dir = tempfile.mkdtemp(prefix="fvg_case_")
    case_file = os.path.join(case_dir, "case.toml")
    with open(case_file, "w") as f:
        f.write('[case]\nname = "Uniform run"\n\n[initial]\nstate = { rho = 1.0, u = 0.0, v = 0.0, p = 1.0 }\n')
    n, U, _ = riemann_ensemble()
    _ENS.update(L=L, case=case_file.encode(), keys=ensemble_keys(), n=n, U=U, off=np.concatenate([[0], np.cumsum(n)]))


def _ens_run(args):
    """Instance i for `steps` steps (or to t_end) in this process: its cell-stage updates."""
    i, steps = args
    L, keys = _ENS["L"], dict(_ENS["keys"], nx=str(int(_ENS["n"][i])))
    ks = [k.encode() for k in keys]
    vs = [str(v).encode() for v in keys.values()]
    h = L.ref_create(_ENS["case"], len(ks), (C.c_char_p * len(ks))(*ks), (C.c_char_p * len(vs))(*vs))
    if not h:
        return 0
    U = np.zeros((int(_ENS["n"][i]), 5))
    U[:, :4] = _ENS["U"][_ENS["off"][i]:_ENS["off"][i + 1]]
    L.ref_set_state(h, U.ctypes.data_as(C.POINTER(C.c_double)))
    dt, sm = C.c_double(), C.c_double()
    t, k = 0.0, 0
    while k < steps and t < 0.15:
        L.ref_step(h, t, 0.15, 0.0, C.byref(dt), C.byref(sm))
        t += dt.value
        k += 1
    L.ref_destroy(h)
    return int(_ENS["n"][i]) * 2 * k


def cpu_ensemble_worker(procs: int, steps: int) -> dict:
    """The reference on the ensemble, process-parallel: every instance for the first `steps` steps in one of `procs`
    single-threaded processes."""
    import multiprocessing as mp
    from fenice_amd.workloads import riemann_ensemble
    ninst = len(riemann_ensemble()[0])
    with mp.get_context("fork").Pool(procs, initializer=_ens_init) as pool:
        pool.map(_ens_run, [(i, 1) for i in range(procs)])   # warm every process (library, instances)
        t0 = time.perf_counter()
        done = sum(pool.map(_ens_run, [(i, steps) for i in range(ninst)], chunksize=1))
        el = time.perf_counter() - t0
    return {"value": done / el / 1e6, "unit": "Mcell-updates/s per RK stage", "cores": procs, "kind": "reference",
            "cpu_model": cpu_model(), "sample": f"the 1024 instances, each for its first {steps} steps (SSP-RK2), in {procs} single-threaded processes "
                      f"({el:.1f} s)"}


def cpu_baseline(n: int, budget_s: float, quadrature: str = "midpoint", workload: str = "tgv", dump: str = ""):
    env = dict(os.environ)
    env.pop("FENICE_NO_SIMD", None)
    if workload == "ensemble":
        env["OMP_NUM_THREADS"] = "1"
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-worker", "--n", str(n), "--cpu-budget", str(budget_s),
                        "--quadrature", quadrature, "--workload", workload] + (["--dump", dump] if dump else []),
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


def measure(U0, n, flux, indicator, weights, steps, warmup):
    """Warm-up steps, then `steps` timed steps of the workload with per-kernel event timing on the context's
    stream: (seconds, steps, kernel times, totals before, totals after, indicator statistics or None)."""
    ind_stats = None
    with make_solver(n, flux, indicator, weights) as s:
        s.upload(U0)
        t = 0.0
        if warmup > 0:
            t, _, _ = s.advance(t, 1e9, max_steps=warmup)
        tot0 = s.conserved_totals()
        s.set_timing(True)
        t0 = time.perf_counter()
        t, k, _ = s.advance(t, 1e9, max_steps=steps)
        el = time.perf_counter() - t0
        times = s.kernel_times()
        tot1 = s.conserved_totals()
        if indicator == "learned":
            last, total, off, audits = s.indicator_stats()
            ind_stats = {"audits": audits, "switched_off": off, "tp_fp_fn_tn_all_audits": total,
                         "missed_troubled_fraction": total[2] / max(1, total[0] + total[2]),
                         "smooth_fraction_of_faces": (total[2] + total[3]) / max(1, sum(total))}
    return el, k, times, tot0, tot1, ind_stats


def compare_with_reference(U0, n, flux, indicator, weights, path, steps) -> dict:
    """The device's state after `steps` steps of the workload from U0 against the reference's (path, .npy)."""
    with make_solver(n, flux, indicator, weights) as s:
        s.upload(U0)
        t = 0.0
        for _ in range(steps):
            t += s.step(t, 1e9).dt
        Ud = s.download()
    Ur = np.load(path)
    same = bool(np.array_equal(Ud.view(np.uint64), Ur.view(np.uint64)))
    d = np.abs(Ud[:, 0] - Ur[:, 0])
    out = {"mesh": n, "steps": steps, "t": t, "bitwise_equal": same, "l1_rel": float(d.sum() / np.abs(Ur[:, 0]).sum()),
           "linf_rel": float(d.max() / np.abs(Ur[:, 0]).max())}
    del Ud, Ur
    return out


def other_modes(U0, n, steps, weights):
    """The other modes on the same workload: FP32 fluxes, and the learned indicator with FP64 and with FP32 fluxes
    (when trained weights exist), each with its density error against the plain FP64 path."""
    from fenice_amd.train_indicator import DEFAULT_OUT
    out = {}
    modes = [("fp32", "fp32", "none")]
    if os.path.exists(weights or DEFAULT_OUT):
        modes += [("fp64+learned", "fp64", "learned"), ("fp32+learned", "fp32", "learned")]
    for name, flux, ind in modes:
        el, k, times, _, _, st = measure(U0, n, flux, ind, weights, steps, 2)
        r = {"value": round(n ** 3 * 3 * k / el / 1e6, 3), "ms_per_step": round(el / k * 1e3, 4), "steps": k,
             "error_vs_fp64": error_vs_fp64(min(n, 128), flux, ind, weights)}
        if st is not None:
            r["indicator"] = st
        out[name] = r
    return out


def side_workload(name: str) -> dict:
    """Throughput of the other configurations of the path (parity cases, not the headline): the 1-D ensemble
    (1024 Riemann problems of 128..8192 cells to t = 0.15, one launch after the autotuner, the largest instances on
    teams of workgroups), the isentropic vortex at
    256^2 under MUSCL-Hancock (CFL 0.7, 200 steps, graph replay), and the Kelvin-Helmholtz layer at 2048^2
    (WENO5-Z + SSP-RK3, 30 steps), the 512^3 Taylor-Green vortex in its three precision modes (with the kinetic energy,
    its dissipation and the density error against FP64 every 100 steps), and the fourth-order flux quadrature."""
    from fenice_amd import fvg
    from fenice_amd.workloads import ensemble_keys, kelvin_helmholtz, kh_keys, riemann_ensemble
    if name == "ensemble":
        n, U, _ = riemann_ensemble()
        with fvg.Ensemble(fvg.config_from_keys(ensemble_keys()), n) as e:
            e.upload(U)
            block, _ = e.tune(0.0075, 400000)
            r = e.run(0.15, 400000)
            plan = e.plan
        stages = 2.0 * float((n.astype(np.float64) * r.steps).sum())
        # beside it, 2-D instances: 256 isentropic vortices on square periodic meshes of 32..256 cells a side, MUSCL
        # (Van Leer) + HLLC + SSP-RK2, CFL 0.5, to t = 1 (the vortex moves by one unit of its 10-wide box)
        from fenice_amd.workloads import vortex_ensemble, vortex_ensemble_keys
        nv, Uv = vortex_ensemble()
        with fvg.Ensemble(fvg.config_from_keys(vortex_ensemble_keys()), nv, ny=nv) as e:
            e.upload(Uv)
            vblock, _ = e.tune(0.05, 100000)
            rv = e.run(1.0, 100000)
            vplan = e.plan
        vstages = 2.0 * float(((nv.astype(np.float64) ** 2) * rv.steps).sum())
        return {"workload": "1024 1-D Riemann problems, 128..8192 cells (4,243,015), MUSCL MC + HLLC + SSP-RK2, CFL 0.5, "
                            "to t = 0.15", "block": block, "kernel_ms": round(r.ms, 2),
                "teams": {"teams": plan[0], "workgroups_in_teams": plan[1], "workgroups": plan[2]},
                "value": round(stages / (r.ms * 1e-3) / 1e6, 1), "unit": "Mcell-updates/s per RK stage",
                "instances_done": int((r.t >= 0.15).sum()), "diverged": int(r.diverged.sum()),
                "steps_max": int(r.steps.max()), "steps_total": int(r.steps.sum()),
                "cpu_baseline": cpu_baseline(16, 20, workload="ensemble"),
                "vortex_instances": {"workload": f"256 isentropic vortices, periodic [0, 10]^2, 32..256 cells a side "
                                                 f"({int((nv.astype(np.int64) ** 2).sum()):,} cells), MUSCL Van Leer + HLLC + "
                                                 "SSP-RK2, CFL 0.5, to t = 1",
                                     "block": vblock, "kernel_ms": round(rv.ms, 2),
                                     "teams": {"teams": vplan[0], "workgroups_in_teams": vplan[1], "workgroups": vplan[2]},
                                     "value": round(vstages / (rv.ms * 1e-3) / 1e6, 1), "unit": "Mcell-updates/s per RK stage",
                                     "instances_done": int((rv.t >= 1.0).sum()), "diverged": int(rv.diverged.sum()),
                                     "steps_max": int(rv.steps.max()), "steps_total": int(rv.steps.sum())}}
    if name == "vortex":
        from fenice_amd.workloads import vortex_keys, vortex_state
        with fvg.Solver(fvg.config_from_keys(vortex_keys(256, cfl="0.7"))) as s:
            s.upload(vortex_state(256))
            s.advance(0.0, 1e9, max_steps=5)
            t0 = time.perf_counter()
            _, k, _ = s.advance(0.0, 1e9, max_steps=200)
            el = time.perf_counter() - t0
        return {"workload": "isentropic vortex 256^2, MUSCL Van Leer + HLLC, MUSCL-Hancock, CFL 0.7, 200 steps "
                            "(graph replay)", "value": round(65536 * k / el / 1e6, 1), "unit": "Mcell-updates/s per RK stage",
                "ms_per_step": round(el / k * 1e3, 4), "steps": k, "cpu_baseline": cpu_baseline(256, 10.0, workload="vortex")}
    if name == "kh":
        # plain FP64 WENO5-Z (the reference's bits), then the learned indicator with FP64 and with FP32 fluxes and plain
        # FP32: each from the initial field, 2 warm-up steps and 30 timed steps at a fixed dt (9e-5, below the CFL step
        # of 1.0e-4) so that every mode reaches the same time; the density's L1 / Linf error against the FP64 run and
        # the audit counts (every 16 steps) are reported. The layer's density jumps are troubled faces from the start.
        from fenice_amd.train_indicator import DEFAULT_OUT, load
        n, dt, steps = 2048, 9e-5, 30
        U0 = kelvin_helmholtz(n)
        params, thr, _, _ = load(DEFAULT_OUT)
        modes, ref, out = [("fp64", False), ("fp64+learned", True), ("fp32", False), ("fp32+learned", True)], None, {}
        for name_m, learned in modes:
            keys = dict(kh_keys(n), flux_precision=name_m[:4])
            if learned:
                keys.update(indicator="learned", indicator_audit="16")
            with fvg.Solver(fvg.config_from_keys(keys)) as s:
                if learned:
                    s.set_indicator(params, thr)
                s.upload(U0)
                t, _, _ = s.advance(0.0, 1e9, max_steps=2, dt_max=dt)
                t0 = time.perf_counter()
                t, k, _ = s.advance(t, 1e9, max_steps=steps, dt_max=dt)
                el = time.perf_counter() - t0
                rho = s.download()[:, 0]
                st = s.indicator_stats() if learned else None
            m = {"value": round(n * n * 3 * k / el / 1e6, 1), "ms_per_step": round(el / k * 1e3, 4), "steps": k, "t": t}
            if ref is None:
                ref = rho
            else:
                d = np.abs(rho - ref)
                m["error_vs_fp64"] = {"l1_rel": float(d.sum() / np.abs(ref).sum()), "linf_rel": float(d.max() / np.abs(ref).max())}
            if st is not None:
                _, total, off, audits = st
                m["indicator"] = {"audits": audits, "switched_off": off, "tp_fp_fn_tn_all_audits": total,
                                  "troubled_fraction": (total[0] + total[2]) / max(1, sum(total)),
                                  "missed_troubled_fraction": total[2] / max(1, total[0] + total[2])}
            out[name_m] = m
        r = out.pop("fp64")
        return {"workload": f"Kelvin-Helmholtz 2048^2, WENO5-Z + HLLC + SSP-RK3, periodic, {steps} steps at dt {dt}",
                "value": r["value"], "unit": "Mcell-updates/s per RK stage", "ms_per_step": r["ms_per_step"], "steps": r["steps"],
                "modes": out, "cpu_baseline": cpu_baseline(n, 15.0, workload="kh")}
    if name == "tgv512":
        # the 512^3 Taylor-Green vortex with FP32 fluxes and the learned indicator beside the plain FP64 path: each mode
        # is timed on 8 steps after 1 warm-up step, then restarted from the initial field for 300 steps at a fixed dt
        # (3e-4, below the CFL step of 3.6e-4); every 100 steps the kinetic energy (volume average) and, against the FP64 run, the
        # density's L1 and Linf errors (relative) are recorded, and the kinetic-energy dissipation rate -dE_k/dt over
        # the last 100 steps
        n = 512
        U0 = tgv_state(n)
        cells = n ** 3
        dt, marks, per = 3e-4, 3, 100
        modes, rho64 = {}, []
        for label, flux, ind in (("fp64", "fp64", "none"), ("fp32", "fp32", "none"), ("fp32+learned", "fp32", "learned")):
            with make_solver(n, flux, ind) as s:
                s.upload(U0)
                s.advance(0.0, 1e9, max_steps=1)
                t0 = time.perf_counter()
                _, k, _ = s.advance(0.0, 1e9, max_steps=8)
                el = time.perf_counter() - t0
                m = {"value": round(cells * 3 * k / el / 1e6, 1), "ms_per_step": round(el / k * 1e3, 3), "steps": k, "history": []}
                print(f"tgv512 {label}: {m['value']} Mcell-updates/s per stage", file=sys.stderr, flush=True)
                s.upload(U0)
                t = 0.0
                ke_prev = float(s.conserved_totals()[1])
                for mark in range(1, marks + 1):
                    for _ in range(per):
                        t += s.step(t, 1e9, dt).dt
                    ke = float(s.conserved_totals()[1])
                    rho = s.download()[:, 0].copy()
                    h = {"step": mark * per, "t": round(t, 6), "kinetic_energy": ke, "dissipation": (ke_prev - ke) / (per * dt)}
                    ke_prev = ke
                    if label == "fp64":
                        rho64.append(rho)
                    else:
                        d = np.abs(rho - rho64[mark - 1])
                        h["l1_rel"] = float(d.sum() / np.abs(rho64[mark - 1]).sum())
                        h["linf_rel"] = float(d.max() / np.abs(rho64[mark - 1]).max())
                    del rho
                    m["history"].append(h)
                    print(f"tgv512 {label}: step {mark * per} {h}", file=sys.stderr, flush=True)
                if ind == "learned":
                    _, total, off, audits = s.indicator_stats()
                    m["indicator"] = {"audits": audits, "switched_off": off, "tp_fp_fn_tn_all_audits": total}
            if label != "fp64":
                last = m["history"][-1]
                m["error_vs_fp64"] = {"l1_rel": last["l1_rel"], "linf_rel": last["linf_rel"], "steps": marks * per, "dt": dt}
            modes[label] = m
        base = modes["fp64"]["value"]
        for m in modes.values():
            m["vs_fp64"] = round(m["value"] / base, 3)
        return {"workload": "3-D Taylor-Green vortex 512^3, Navier-Stokes Re 1600 Pr 0.71 Mach 0.1, WENO5-Z + HLLC + SSP-RK3, "
                            "CFL 0.5, periodic", "unit": "Mcell-updates/s per RK stage", "modes": modes}
    if name == "fourth":
        # the 256^3 Taylor-Green vortex with the fourth-order flux quadrature beside the midpoint one (FP64 both): 2
        # warm-up steps, 10 timed steps with per-kernel event timing; the reference's own code on the host at 96^3
        n = 256
        U0 = tgv_state(n)
        cells = n ** 3
        nf = 3 * (n + 1) * n * n
        modes = {}
        for quad in ("midpoint", "fourth"):
            with fvg.Solver(fvg.config_from_keys(run_keys(n, "fp64", quad))) as s:
                s.upload(U0)
                s.advance(0.0, 1e9, max_steps=2)
                s.set_timing(True)
                t0 = time.perf_counter()
                _, k, _ = s.advance(0.0, 1e9, max_steps=10)
                el = time.perf_counter() - t0
                times = s.kernel_times()
            m = {"value": round(cells * 3 * k / el / 1e6, 1), "ms_per_step": round(el / k * 1e3, 3), "steps": k,
                 "kernels": {kn: {"avg_ms": round(v[0] / max(1, v[1]), 4), "launches": v[1]} for kn, v in times.items()}}
            if quad == "fourth":
                # algorithmic HBM bytes per face of each pass (FP64, NV = 5): states write two states and a flag
                # (81 B) and read the padded conservative stencil once per cell (40 B); points read the states and
                # the flag (81 B) and write the flux and a flag (41 B); fluxes read the flux and flag (41 B) and write
                # the convective (40 B) and viscous (32 B) fluxes (the viscous inputs, W and G, are not counted)
                per_face = {"k4_states": 121.0, "k4_points": 122.0, "k4_fluxes": 113.0}
                for kn, b in per_face.items():
                    if kn in times and times[kn][1]:
                        avg_s = times[kn][0] / times[kn][1] * 1e-3
                        m["kernels"][kn]["algorithmic_GBs"] = round(b * nf / 3 / avg_s / 1e9, 1)
            modes[quad] = m
        modes["fourth"]["vs_midpoint"] = round(modes["fourth"]["value"] / modes["midpoint"]["value"], 3)
        return {"workload": "3-D Taylor-Green vortex 256^3, Navier-Stokes Re 1600 Pr 0.71 Mach 0.1, WENO5-Z + HLLC + SSP-RK3, "
                            "CFL 0.5, periodic, FP64", "unit": "Mcell-updates/s per RK stage", "modes": modes,
                "cpu_baseline_fourth": cpu_baseline(96, 15.0, "fourth")}
    raise ValueError(name)


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
    ap.add_argument("--indicator", default="none", choices=["none", "learned"], help="troubled-face indicator (WENO5-Z, either flux precision)")
    ap.add_argument("--passes", default="default", choices=["default", "fused", "separate"],
                    help="kernel sequence of the residual: fused face passes (no face fluxes in HBM) or separate face and update passes")
    ap.add_argument("--weights", default="", help="trained indicator weights (default fenice_amd/indicator_weights.npz)")
    ap.add_argument("--no-modes", action="store_true", help="skip the FP32 / learned-indicator lines of the FP64 run")
    ap.add_argument("--quadrature", default="midpoint", choices=["midpoint", "fourth"], help="flux quadrature of the CPU worker")
    ap.add_argument("--workload", default="tgv", choices=["tgv", "vortex", "kh", "ensemble"], help="workload of the CPU worker")
    ap.add_argument("--dump", default="", help="CPU worker: write its final state here (.npy)")
    ap.add_argument("--no-side", action="store_true", help="skip the vortex and Kelvin-Helmholtz lines of the default run")
    ap.add_argument("--side", default="", choices=["", "ensemble", "vortex", "kh", "tgv512", "fourth"],
                    help="measure one of the other configurations instead (prints its own JSON line)")
    a = ap.parse_args()
    global PASSES
    PASSES = a.passes
    if a.cpu_worker:
        if a.workload == "ensemble":   # --n: processes, --cpu-budget: steps per instance
            print(json.dumps(cpu_ensemble_worker(a.n, int(a.cpu_budget))))
        else:
            print(json.dumps(cpu_baseline_worker(a.n, a.cpu_budget, a.quadrature, a.workload, a.dump)))
        return
    if a.side:
        print(json.dumps(side_workload(a.side)), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: the library has no multi-GPU path (asked for {a.gpus} GPUs); the mesh would "
              "be cut into z-slabs, one per GPU, with halo exchanges between neighbours", file=sys.stderr)
        sys.exit(2)

    n = a.n
    U0 = tgv_state(n)
    cells = n ** 3
    el, steps, times, tot0, tot1, ind_stats = measure(U0, n, a.flux, a.indicator, a.weights, a.steps, a.warmup)
    stages = 3 * steps
    value = cells * stages / el / 1e6
    # the stage as the unit: algorithmic bytes of one stage over the event-measured time of its kernels
    fam = {k: v for k, v in times.items() if v[1] > 0}
    stage_ms = sum(ms for ms, _ in fam.values()) / stages
    achieved = BYTES_PER_CELL_STAGE_RK3_3D * cells / (stage_ms * 1e-3) / 1e9
    dom = max(fam.items(), key=lambda kv: kv[1][0])
    traffic = MEASURED_STAGE_TRAFFIC.get((a.flux, n)) if a.indicator == "none" else None
    # the dominant kernel's VALU issue: its instructions per wave (counter run) x waves / (SIMDs x clock / 4 x time)
    valu = None
    vt, vsrc = VALU_PER_WAVE.get(a.flux, ({}, ""))
    if n == 256 and a.indicator == "none" and dom[0] in vt:
        waves = ((n + 1) * n * n + 63) // 64 if dom[0].startswith("k_faces") else cells // 64
        avg_s = dom[1][0] / dom[1][1] * 1e-3
        issue = vt[dom[0]] * waves * 4.0 / (SIMDS * CLOCK_HZ * avg_s)
        valu = {"kernel": dom[0], "instr_per_wave": vt[dom[0]], "waves": waves, "avg_ms": round(avg_s * 1e3, 4),
                "issue_frac": round(issue, 4), "instr_per_wave_source": vsrc}
    out = {
        "metric": "Mcell-updates/s per RK stage, % HBM peak, at stated L1 error vs FP64 oracle",
        "value": round(value, 3), "unit": "Mcell-updates/s per RK stage", "n_gpus": 1, "steps": steps, "warmup": a.warmup,
        "ms_per_step": round(el / steps * 1e3, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f64" if a.flux == "fp64" else "f32 fluxes, f64 state") + (", learned indicator (bf16 MFMA)" if a.indicator == "learned" else ""),
        "data": "synthetic (analytic Taylor-Green initial field)",
        "config": {"workload": f"3-D Taylor-Green vortex {n}^3, Navier-Stokes Re 1600 Pr 0.71 Mach 0.1, WENO5-Z + HLLC + SSP-RK3, "
                               "CFL 0.5, periodic", "cells": cells, "stages_per_step": 3},
        "l1_error_vs_fp64_oracle": None,
        # the bound is the vector ALU (the face kernels' FP64 arithmetic); achieved / peak / frac are the HBM figures
        # the metric asks for (algorithmic bytes of a stage over the stage's kernel time), valu the issue rate
        "roofline": {"bound": "valu" if a.flux == "fp64" else "valu/hbm",
                     "kernel": "one RK stage (" + ", ".join(k for k in fam if k != "k_control") + ")",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": traffic[0] if traffic else None, "traffic_source": traffic[1] if traffic else None,
                     "bytes_per_cell_stage": round(BYTES_PER_CELL_STAGE_RK3_3D, 2), "valu": valu},
        "kernels": {k: {"ms_total": round(ms, 3), "launches": cnt, "avg_ms": round(ms / cnt, 4)} for k, (ms, cnt) in fam.items()},
        "dominant_kernel": dom[0],
        "conservation": {"mass_rel_drift": abs(tot1[0] - tot0[0]) / abs(tot0[0]), "energy_rel_drift": abs(tot1[2] - tot0[2]) / abs(tot0[2])},
    }
    if ind_stats is not None:
        out["indicator"] = ind_stats
    if a.flux != "fp64" or a.indicator != "none":
        out["error_vs_fp64"] = error_vs_fp64(min(n, 128), a.flux, a.indicator, a.weights)
        out["l1_error_vs_fp64_oracle"] = out["error_vs_fp64"]["l1_rel"]
    if a.flux == "fp64" and a.indicator == "none" and not a.no_modes:
        out["other_modes"] = other_modes(U0, n, max(10, a.steps // 3), a.weights)
    if not a.no_cpu:
        import tempfile
        dump = os.path.join(tempfile.mkdtemp(prefix="fvg_bench_"), "ref_state.npy") if (a.cpu_n or n) == n else ""
        cb = cpu_baseline(a.cpu_n or n, a.cpu_budget, dump=dump)
        if dump and cb.get("state_steps") and os.path.exists(dump):
            # the reference's state after its sample's steps from the same initial field, against the device's after
            # as many steps: the FP64 path must give the same bits (L1 / Linf of the density, relative, otherwise)
            out["parity_vs_reference"] = compare_with_reference(U0, n, a.flux, a.indicator, a.weights, dump, cb["state_steps"])
            if a.flux == "fp64" and a.indicator == "none":
                out["l1_error_vs_fp64_oracle"] = out["parity_vs_reference"]["l1_rel"]
            os.remove(dump)
        cb.pop("state_steps", None)
        cb.pop("state_t", None)
        out["cpu_baseline"] = cb
    if a.flux == "fp64" and a.indicator == "none" and not a.no_side:
        out["side"] = {name: side_workload(name) for name in ("vortex", "kh")}
    print(json.dumps(out), flush=True)
    if a.calibrate_copy:
        copy_calibration()


if __name__ == "__main__":
    main()
