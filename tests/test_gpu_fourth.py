"""flux_quadrature = fourth on the device against the oracle (which equals the reference bit for bit): residual,
wave speed and whole steps over the fourth-order configurations, graph replays at a larger mesh, and a 128^3 run
that must be deterministic and conservative."""
import sys

import numpy as np
import pytest

from cases import base_keys, fourth_matrix, shock_state, smooth_state, state_for
from fenice_amd import fvg

pytestmark = pytest.mark.gpu
FOURTH = fourth_matrix()


def same_bits(a, b, two_d=False):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    if two_d:
        return np.array_equal(a[:, :4].view(np.uint64), b[:, :4].view(np.uint64))
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,keys,kind", FOURTH, ids=[m[0] for m in FOURTH])
def test_fourth_residual_and_steps_match_oracle(oracle, name, keys, kind):
    say(f"fourth parity {name}")
    cfg = fvg.config_from_keys(keys)
    two_d = cfg.nz == 1
    U = state_for(kind, keys)
    with fvg.Solver(cfg) as s:
        s.upload(U)
        Rd, Sd = s.residual()
        Ro, So = oracle.residual(cfg, U)
        assert Sd == So
        assert same_bits(Rd, Ro, two_d)
        Uo = U.copy()
        t, lag = 0.0, 0.0
        for _ in range(3):
            r = s.step(t, 10.0)
            dto, so, lag, ok = oracle.step(cfg, Uo, t, 10.0, 0.0, lag)
            assert ok and r.dt == dto and r.s_max == so
            t += r.dt
        assert same_bits(s.download(), Uo, two_d)


@pytest.mark.parametrize("case", ["3d-periodic-visc", "3d-walls-shock", "2d-walls-shock"])
def test_fourth_graph_runs_match_oracle(oracle, case):
    # advance() replays captured graphs of four steps; several calls with eager work between them
    say("fourth graph run", case)
    if case == "3d-periodic-visc":
        n = (26, 22, 20)
        keys = base_keys(*n, reconstruction="weno5", use_viscous="true", mu="0.01", pr="0.71", flux_quadrature="fourth")
        U = smooth_state(*n, 51, mach=0.3)
    elif case == "3d-walls-shock":
        n = (24, 20, 18)
        keys = base_keys(*n, periodic=False, reconstruction="weno5", boundary_order="2", flux_quadrature="fourth",
                         bc_xmin="transmissive", bc_xmax="wall", bc_ymin="no_slip", bc_ymax="wall", bc_zmin="wall",
                         bc_zmax="transmissive")
        U = shock_state(*n, 52)
    else:
        n = (70, 46, 1)
        keys = base_keys(*n[:2], periodic=False, reconstruction="weno3", boundary_order="2", flux_quadrature="fourth",
                         bc_xmin="transmissive", bc_xmax="wall", bc_ymin="no_slip", bc_ymax="wall")
        U = shock_state(*n, 53)
    cfg = fvg.config_from_keys(keys)
    Uo = U.copy()
    t = 0.0
    with fvg.Solver(cfg) as s:
        s.upload(U)
        for chunk in (5, 3):
            td, nd, _ = s.advance(t, 100.0, max_steps=chunk)
            for _ in range(chunk):
                dt, _, _, ok = oracle.step(cfg, Uo, t, 100.0, 0.0, 0.0)
                assert ok
                t += dt
            assert nd == chunk and td == t
            assert same_bits(s.download(), Uo, cfg.nz == 1)
            s.residual()   # eager work between the replays


def test_fourth_large_run_is_deterministic_and_conservative():
    say("fourth 128^3 determinism")
    n = 128
    keys = base_keys(n, n, n, reconstruction="weno5", use_viscous="true", mu="6.25e-4", pr="0.71", flux_quadrature="fourth")
    cfg = fvg.config_from_keys(keys)
    U = shock_state(n, n, n, 6)
    out = []
    for _ in range(2):
        with fvg.Solver(cfg) as s:
            s.upload(U)
            tot0 = s.conserved_totals()
            s.advance(0.0, 100.0, max_steps=3)
            out.append((s.download(), s.conserved_totals()))
    assert same_bits(out[0][0], out[1][0])
    tot1 = out[0][1]
    assert abs(tot1[0] - tot0[0]) <= 1e-13 * abs(tot0[0])
    assert abs(tot1[2] - tot0[2]) <= 1e-13 * abs(tot0[2])


def test_fourth_fp32_is_refused():
    keys = base_keys(8, 8, reconstruction="weno5", flux_quadrature="fourth", flux_precision="fp32")
    with pytest.raises(fvg.Unsupported):
        fvg.Solver(fvg.config_from_keys(keys))


@pytest.mark.parametrize("dims", [(16, 12, 1), (8, 9, 7)], ids=["2d", "3d"])
def test_fourth_quotients_outside_the_fast_range(oracle, dims):
    # subnormal transverse momenta: k_cells finds padded conservative values below 2^-900 and the conserved WENO5-Z
    # of k4_states takes the IEEE division by 6 (Schemes.cpp:343-345); the residual and steps keep the reference's
    # bits, and a field in range afterwards takes the three-operation quotients with the same bits
    from test_gpu_parity import tiny_velocity_state
    nx, ny, nz = dims
    say(f"fourth quotients outside the fast range {dims}")
    keys = base_keys(nx, ny, nz, reconstruction="weno5", flux_quadrature="fourth")
    cfg = fvg.config_from_keys(keys)
    two_d = nz == 1
    U = tiny_velocity_state(nx, ny, nz, 7)
    with fvg.Solver(cfg) as s:
        s.upload(U)
        Rd, Sd = s.residual()
        Ro, So = oracle.residual(cfg, U)
        assert Sd == So and same_bits(Rd, Ro, two_d)
        Uo = U.copy()
        t, lag = 0.0, 0.0
        for _ in range(3):
            r = s.step(t, 10.0)
            dto, so, lag, ok = oracle.step(cfg, Uo, t, 10.0, 0.0, lag)
            assert ok and r.dt == dto and r.s_max == so
            t += r.dt
        assert same_bits(s.download(), Uo, two_d)
        U2 = smooth_state(nx, ny, nz, 8)
        s.upload(U2)
        Rd, Sd = s.residual()
        Ro, So = oracle.residual(cfg, U2)
        assert Sd == So and same_bits(Rd, Ro, two_d)
