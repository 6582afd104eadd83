"""The oracle against the reference itself (built from its sources): residual, wave speed, the MUSCL-Hancock
predictor and whole steps, bit for bit, over the configuration matrix."""
import numpy as np
import pytest

from cases import VORTEX_DIVERGES_AT, VORTEX_HANCOCK_256, matrix, state_for, vortex_state
from fenice_amd.fvg import config_from_keys

MATRIX = matrix()


@pytest.mark.parametrize("name,keys,kind", MATRIX, ids=[m[0] for m in MATRIX])
def test_oracle_equals_reference(oracle, ref, name, keys, kind):
    case = "Taylor-Green Vortex 3D" if int(keys["nz"]) > 1 else "Isentropic Vortex"
    run = ref.create(case, **keys)
    cfg = config_from_keys(keys)
    U = state_for(kind, keys)
    Rr, Sr = run.residual(U)
    Ro, So = oracle.residual(cfg, U)
    assert np.array_equal(Rr.view(np.uint64), Ro.view(np.uint64)) and Sr == So
    if keys["time_scheme"] == "MUSCL-Hancock":
        Rr, Sr = run.residual(U, 0.002)
        Ro, So = oracle.residual(cfg, U, 0.002)
        assert np.array_equal(Rr.view(np.uint64), Ro.view(np.uint64)) and Sr == So
    run.set_state(U)
    Uo = U.copy()
    t, lag = 0.0, 0.0
    for _ in range(3):
        dtr, sr = run.step(t, 10.0)
        dto, so, lag, ok = oracle.step(cfg, Uo, t, 10.0, 0.0, lag)
        assert ok and dtr == dto and sr == so
        t += dtr
    assert np.array_equal(run.get_state().view(np.uint64), Uo.view(np.uint64))


def test_vortex_hancock_cfl08_diverges_at_65(oracle, ref):
    # every step of the unstable MUSCL-Hancock vortex run is bit-identical, and both sides reach the same step at
    # which t + dt == t (the reference's check_progress, SimulationInternal.hpp:259-267)
    keys = VORTEX_HANCOCK_256
    run = ref.create("Isentropic Vortex", **keys)
    cfg = config_from_keys(keys)
    U = vortex_state(256)
    run.set_state(U)
    t, lag = 0.0, 0.0
    for step in range(1, VORTEX_DIVERGES_AT + 1):
        dtr, sr = run.step(t, 1e9)
        dto, so, lag, ok = oracle.step(cfg, U, t, 1e9, 0.0, lag)
        assert dtr == dto and sr == so, step
        assert np.array_equal(run.get_state().view(np.uint64), U.view(np.uint64)), step
        progressed = np.isfinite(sr) and np.isfinite(dtr) and t + dtr > t
        assert ok == progressed and progressed == (step < VORTEX_DIVERGES_AT), step
        t += dtr


@pytest.mark.parametrize("n", [1, 2, 3, 5, 40])
@pytest.mark.parametrize("recon,bo", [("muscl", "1"), ("muscl", "2"), ("weno5", "2"), ("none", "2")])
def test_one_row_runs_equal_reference(oracle, ref, n, recon, bo):
    # the N x 1 mapping of a 1-D problem (periodic y, one row), down to meshes smaller than the halo depth
    from fenice_amd.workloads import ensemble_keys, riemann_ensemble, riemann_states
    _, _, params = riemann_ensemble(3, seed=13, nmin=1, nmax=2)
    keys = dict(ensemble_keys(), nx=str(n), reconstruction=recon, boundary_order=bo)
    if recon == "weno5":
        keys["time_scheme"] = "SSP Runge-Kutta 3"
    U4 = riemann_states([n], params[1:2])
    U = np.zeros((n, 5))
    U[:, :4] = U4
    run = ref.create("Sod Shock Tube", **keys)
    run.set_state(U)
    cfg = config_from_keys(keys)
    t, lag, steps = 0.0, 0.0, 0
    while t < 0.15 and steps < 12:
        dtr, sr = run.step(t, 0.15)
        dto, so, lag, ok = oracle.step(cfg, U, t, 0.15, 0.0, lag)
        assert ok and dtr == dto and sr == so
        t += dtr
        steps += 1
    assert np.array_equal(run.get_state().view(np.uint64), U.view(np.uint64))
    assert np.all(U[:, 2] == 0.0) and not np.signbit(U[:, 2]).any()


@pytest.mark.parametrize("dims,kind", [((12, 10, 1), "shock"), ((8, 9, 7), "smooth"), ((5, 3, 3), "shock")])
def test_volume_averages_equal_reference(oracle, ref, dims, kind):
    # fvo_volume_averages (the checker of the device's conserved totals and monitor row) against the reference's
    # own Simulation::volume_averages, bit for bit
    nx, ny, nz = dims
    from cases import base_keys
    keys = base_keys(nx, ny, nz)
    case = "Taylor-Green Vortex 3D" if nz > 1 else "Isentropic Vortex"
    run = ref.create(case, **keys)
    cfg = config_from_keys(keys)
    U = state_for(kind, keys, seed=23)
    a = oracle.volume_averages(cfg, U)
    b = run.volume_averages(U)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
