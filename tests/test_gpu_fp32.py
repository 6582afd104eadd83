"""The FP32-flux mode: reconstruction, Riemann solver and viscous fluxes in FP32; state, stage combination, dt
reduction and conservation sums in FP64. Its error is stated against the FP64 path (= the oracle's bits)."""
import numpy as np
import pytest

from cases import base_keys, smooth_state, shock_state
from fenice_amd import fvg

pytestmark = pytest.mark.gpu


def say(what):
    print(f"[fp32] {what}", flush=True)


def run(keys, U, steps, dt):
    cfg = fvg.config_from_keys(keys)
    with fvg.Solver(cfg) as s:
        s.upload(U)
        t = 0.0
        for _ in range(steps):
            t += s.step(t, 1e9, dt).dt
        return s.download(), s.conserved_totals()


@pytest.mark.parametrize("rc", ["weno5", "weno3", "muscl"])
def test_fp32_residual_close_to_fp64(oracle, rc):
    say(f"residual {rc}")
    n = 24
    keys = base_keys(n, n, n, reconstruction=rc, use_viscous="true", mu="6.25e-4", pr="0.71")
    U = smooth_state(n, n, n, 3, mach=0.3)
    R64, S64 = oracle.residual(fvg.config_from_keys(keys), U)
    cfg32 = fvg.config_from_keys(dict(keys, flux_precision="fp32"))
    with fvg.Solver(cfg32) as s:
        s.upload(U)
        R32, S32 = s.residual()
    scale = np.abs(R64).max(axis=0)
    err = np.abs(R32 - R64).max(axis=0) / scale
    assert np.all(err < 2e-4), err
    assert abs(S32 - S64) <= 1e-5 * S64


def test_fp32_run_error_and_conservation():
    # 3-D Taylor-Green-like flow at Mach 0.1 (the benchmark's regime), 20 steps at a fixed dt
    say("tgv-like 32^3 run")
    n = 32
    keys = base_keys(n, n, n, reconstruction="weno5", use_viscous="true", mu="6.25e-4", pr="0.71", riemann_solver="HLLC")
    U = smooth_state(n, n, n, 8, mach=0.1)
    U64, tot64 = run(keys, U, 20, 2e-3)
    U32, tot32 = run(dict(keys, flux_precision="fp32"), U, 20, 2e-3)
    l1 = np.abs(U32[:, 0] - U64[:, 0]).sum() / np.abs(U64[:, 0]).sum()
    assert l1 < 1e-6, l1
    with fvg.Solver(fvg.config_from_keys(keys)) as s:
        s.upload(U)
        tot0 = s.conserved_totals()
    # the FP64 divergence of one flux per face telescopes: mass and energy stay at round-off
    assert abs(tot32[0] - tot0[0]) <= 1e-13 * tot0[0]
    assert abs(tot32[2] - tot0[2]) <= 1e-13 * abs(tot0[2])


def test_fp32_shock_2d_stays_close():
    say("2-d shock run")
    keys = base_keys(64, 48, reconstruction="weno5", periodic=False, bc_xmin="transmissive", bc_xmax="transmissive",
                     bc_ymin="wall", bc_ymax="wall")
    U = shock_state(64, 48, 1, 4)
    U64, _ = run(keys, U, 10, 1e-3)
    U32, _ = run(dict(keys, flux_precision="fp32"), U, 10, 1e-3)
    l1 = np.abs(U32[:, 0] - U64[:, 0]).sum() / np.abs(U64[:, 0]).sum()
    assert l1 < 1e-5, l1


@pytest.mark.parametrize("dims", [(48, 40, 1), (20, 18, 16)])
def test_fp32_inflow_and_nonreflecting_sides_stay_close(dims):
    # a mass-flow inlet and a non-reflecting outlet (their side states in FP64, read by the FP32 faces), with
    # boundary_order 2 ghost slabs: 10 steps at a fixed dt stay at FP32 distance from the FP64 run
    say(f"fp32 inflow / non-reflecting {dims}")
    io = dict(bc_xmin="massflow_inlet", bc_xmax="nonreflecting", bc_ymin="wall", bc_ymax="no_slip", u_in="0.5", v_in="0.1",
              rho_in="1.2", p_out="0.9", boundary_order="2")
    if dims[2] > 1:
        io.update(bc_zmin="periodic", bc_zmax="periodic")
    keys = base_keys(*dims, periodic=False, reconstruction="weno5", **io)
    U = smooth_state(*dims, 9, mach=0.3)
    U64, _ = run(keys, U, 10, 1e-3)
    U32, _ = run(dict(keys, flux_precision="fp32"), U, 10, 1e-3)
    l1 = np.abs(U32[:, 0] - U64[:, 0]).sum() / np.abs(U64[:, 0]).sum()
    assert l1 < 1e-6, l1


def test_fp32_rejects_hancock_predictor():
    say("hancock rejected")
    keys = base_keys(16, 12, reconstruction="muscl", time_scheme="SSP Runge-Kutta 2", flux_precision="fp32")
    with fvg.Solver(fvg.config_from_keys(keys)) as s:
        s.upload(smooth_state(16, 12, 1, 1))
        with pytest.raises(fvg.Unsupported):
            s.residual(0.01)
    with pytest.raises(fvg.Unsupported):
        fvg.Solver(fvg.config_from_keys(dict(keys, time_scheme="MUSCL-Hancock")))
