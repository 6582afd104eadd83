"""The device path against the oracle, bit for bit, and against the reference's stored states.

FP64 parity bar: identical bits (the device code keeps the reference's operation order with no fused
multiply-add and IEEE division / square root). 2-D fields carry no rho*w plane on the device; its slot is +0 on
download and is compared by value."""
import sys

import numpy as np
import pytest

from cases import (VORTEX_DIVERGES_AT, VORTEX_HANCOCK_256, base_keys, golden_names, load_golden, matrix, shock_state, smooth_state,
                   state_for, vortex_state)
from fenice_amd import fvg
from fenice_amd.config import Config, EOS_STIFFENED, HLLC, LOWMACH_HLLC
from riemann_pool import HASHES, fingerprint, pool

pytestmark = pytest.mark.gpu
MATRIX = matrix()


def same_bits(a, b, two_d=False):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    if two_d:
        return np.array_equal(a[:, :4].view(np.uint64), b[:, :4].view(np.uint64)) and np.array_equal(a[:, 4], b[:, 4])
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind", ["ideal", "stiffened"])
@pytest.mark.parametrize("solver", ["hllc", "lowmach"])
def test_device_riemann_fingerprint(kind, solver):
    say(f"riemann fingerprint {kind} {solver}")
    UL, UR, ax, gamma, p_inf, stiff = pool(kind)
    cfg = Config.make(gamma=gamma, p_inf=p_inf, eos_type=EOS_STIFFENED if stiff else 0, shock_blend=0.0,
                      riemann=LOWMACH_HLLC if solver == "lowmach" else HLLC)
    F, S = fvg.riemann_batch(cfg, UL, UR, ax)
    assert fingerprint(F, S) == HASHES[(kind, solver)]


@pytest.mark.parametrize("name,keys,kind", MATRIX, ids=[m[0] for m in MATRIX])
def test_residual_and_steps_match_oracle(oracle, name, keys, kind):
    say(f"parity {name}")
    cfg = fvg.config_from_keys(keys)
    two_d = cfg.nz == 1
    U = state_for(kind, keys)
    with fvg.Solver(cfg) as s:
        s.upload(U)
        Rd, Sd = s.residual()
        Ro, So = oracle.residual(cfg, U)
        assert Sd == So
        assert same_bits(Rd, Ro, two_d)
        if keys["time_scheme"] == "MUSCL-Hancock":
            Rd, Sd = s.residual(0.002)
            Ro, So = oracle.residual(cfg, U, 0.002)
            assert Sd == So and same_bits(Rd, Ro, two_d)
        Uo = U.copy()
        t, lag = 0.0, 0.0
        for _ in range(3):
            r = s.step(t, 10.0)
            dto, so, lag, ok = oracle.step(cfg, Uo, t, 10.0, 0.0, lag)
            assert ok and r.dt == dto and r.s_max == so
            t += r.dt
        assert same_bits(s.download(), Uo, two_d)


@pytest.mark.parametrize("name", golden_names())
def test_device_reproduces_reference_state(oracle, name):
    say(f"golden {name}")
    g = load_golden(name)
    cfg = fvg.config_from_keys(g["keys"])
    t_end = float(g["keys"]["t_end"])
    with fvg.Solver(cfg) as s:
        s.upload(g["U0"])
        t, steps, _ = s.advance(0.0, t_end)
        U = s.download()
    assert steps == g["step"]
    Uref = g["U_ref"]
    scale = np.maximum(np.max(np.abs(Uref), axis=0), 1e-30)
    assert float(np.max(np.abs(U[:, : Uref.shape[1]] - Uref) / scale)) < 1e-8
    Uo = g["U0"].copy()
    oracle.run(cfg, Uo, t_end)
    assert same_bits(U, Uo, cfg.nz == 1)


def test_advance_equals_single_steps():
    say("advance vs step")
    keys = base_keys(16, 12, 10, reconstruction="weno5", use_viscous="true", mu="0.01")
    cfg = fvg.config_from_keys(keys)
    U = smooth_state(16, 12, 10, 4)
    with fvg.Solver(cfg) as a, fvg.Solver(cfg) as b:
        a.upload(U)
        b.upload(U)
        t, n, _ = a.advance(0.0, 1.0, max_steps=5)
        tb = 0.0
        for _ in range(5):
            tb += b.step(tb, 1.0).dt
        assert n == 5 and t == tb
        assert same_bits(a.download(), b.download())


def test_totals_match_volume_averages(oracle):
    say("totals")
    keys = base_keys(20, 18, 16, reconstruction="weno5")
    cfg = fvg.config_from_keys(keys)
    U = smooth_state(20, 18, 16, 11)
    with fvg.Solver(cfg) as s:
        s.upload(U)
        tot = s.conserved_totals()
    ref = oracle.volume_averages(cfg, U)
    assert np.allclose(tot, ref, rtol=1e-13, atol=0)


def test_divergence_is_reported(oracle):
    # a cell with infinite energy gives an infinite wave speed, a zero time step and the reference's
    # check_progress error (a NaN face alone would not: the reference's maximum drops NaN); forced from the host,
    # every kernel still exits normally
    say("divergence")
    keys = base_keys(12, 10, reconstruction="muscl")
    cfg = fvg.config_from_keys(keys)
    U = smooth_state(12, 10, 1, 2)
    U[17] = [1.0, 0.0, 0.0, np.inf, 0.0]
    with fvg.Solver(cfg) as s:
        s.upload(U)
        with pytest.raises(fvg.Diverged):
            s.step(0.0, 1.0)
    _, _, _, ok = oracle.step(cfg, U.copy(), 0.0, 1.0)
    assert not ok


def test_bench_shape_step_matches_oracle(oracle):
    # the benchmark's scheme (3-D Taylor-Green, WENO5-Z + HLLC + SSP-RK3, Navier-Stokes, Pr 0.71) at 48^3, two steps
    say("tgv 48^3 parity")
    n = 48
    keys = base_keys(n, n, n, reconstruction="weno5", use_viscous="true", mu="6.25e-4", pr="0.71", cfl="0.5")
    cfg = fvg.config_from_keys(keys)
    U = smooth_state(n, n, n, 21, mach=0.1)
    with fvg.Solver(cfg) as s:
        s.upload(U)
        t, steps, _ = s.advance(0.0, 100.0, max_steps=2)
        Ud = s.download()
    Uo = U.copy()
    oracle.run(cfg, Uo, 100.0, max_steps=2)
    assert same_bits(Ud, Uo)


def test_large_run_is_deterministic_and_conservative():
    # 128^3 periodic: two runs give the same bits (no race in reductions), mass and energy are conserved to round-off
    say("128^3 determinism")
    n = 128
    keys = base_keys(n, n, n, reconstruction="weno5", use_viscous="true", mu="6.25e-4", pr="0.71")
    cfg = fvg.config_from_keys(keys)
    U = smooth_state(n, n, n, 5, mach=0.2)
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


def test_vortex_hancock_cfl08_per_step_and_divergence(oracle):
    # the isentropic vortex at 256^2, MUSCL-Hancock at CFL 0.8: bit-identical after each of steps 1-64, and the
    # device reports the divergence at step 65 exactly where the oracle (and the reference) do
    say("vortex hancock cfl 0.8")
    keys = VORTEX_HANCOCK_256
    cfg = fvg.config_from_keys(keys)
    U = vortex_state(256)
    Uo = U.copy()
    t, lag = 0.0, 0.0
    with fvg.Solver(cfg) as s:
        s.upload(U)
        for step in range(1, VORTEX_DIVERGES_AT):
            r = s.step(t, 1e9)
            dto, so, lag, ok = oracle.step(cfg, Uo, t, 1e9, 0.0, lag)
            assert ok and r.dt == dto and r.s_max == so, step
            assert same_bits(s.download(), Uo), step
            t += r.dt
        assert s.lagged_speed == lag
        with pytest.raises(fvg.Diverged):
            s.step(t, 1e9)
    assert not oracle.step(cfg, Uo, t, 1e9, 0.0, lag)[3]


@pytest.mark.parametrize("case", ["3d-weno5-visc", "2d-hancock", "2d-inflow-nrbc", "3d-inflow-nrbc", "3d-velin-outlet-z"])
def test_graph_replays_match_eager_steps(oracle, case):
    # advance() replays a captured graph of four steps; checked after each of several calls (with eager work and
    # host synchronisation between the replays) against the kernel-by-kernel launches and the oracle
    say("graph replays", case)
    if case == "3d-weno5-visc":
        keys = base_keys(18, 16, 14, reconstruction="weno5", use_viscous="true", mu="0.01", pr="0.71")
        U = smooth_state(18, 16, 14, 31, mach=0.2)
    elif case == "2d-hancock":
        keys = base_keys(40, 36, time_scheme="MUSCL-Hancock")
        U = smooth_state(40, 36, 1, 32, mach=0.3)
    elif case == "3d-velin-outlet-z":
        keys = base_keys(18, 16, 14, periodic=False, reconstruction="weno5", boundary_order="2", bc_xmin="periodic",
                         bc_xmax="periodic", bc_ymin="wall", bc_ymax="no_slip", bc_zmin="vel_inlet", bc_zmax="outlet", u_in="0.4",
                         v_in="0.1", p_out="0.9")
        U = smooth_state(18, 16, 14, 35, mach=0.3)
    else:
        # the mass-flow target and the non-reflecting ghosts come from kernels of their own inside the graph
        io = dict(bc_xmin="massflow_inlet", bc_xmax="nonreflecting", bc_ymin="wall", bc_ymax="no_slip", u_in="0.5", v_in="0.1",
                  rho_in="1.2", p_out="0.9", boundary_order="2")
        if case == "2d-inflow-nrbc":
            keys = base_keys(40, 36, periodic=False, reconstruction="weno5", **io)
            U = smooth_state(40, 36, 1, 33, mach=0.3)
        else:
            keys = base_keys(18, 16, 14, periodic=False, use_viscous="true", mu="0.01", bc_zmin="periodic", bc_zmax="periodic", **io)
            U = smooth_state(18, 16, 14, 34, mach=0.3)
    cfg = fvg.config_from_keys(keys)
    a, b = fvg.Solver(cfg), fvg.Solver(cfg)
    b.set_graphs(False)
    Uo = U.copy()
    a.upload(U)
    b.upload(U)
    t, lag = 0.0, 0.0
    for k, chunk in enumerate((5, 7, 1, 9)):
        if case == "2d-hancock" and k == 2:   # a reset lagged speed brings the seed residual back (eager, outside the graph)
            a.lagged_speed = 0.0
            b.lagged_speed = 0.0
            lag = 0.0
        ta, na, _ = a.advance(t, 100.0, max_steps=chunk)
        tb, nb, _ = b.advance(t, 100.0, max_steps=chunk)
        for _ in range(chunk):
            dt, _, lag, ok = oracle.step(cfg, Uo, t, 100.0, 0.0, lag)
            assert ok
            t += dt
        Ua = a.download()
        assert na == nb == chunk and ta == tb == t
        assert same_bits(Ua, b.download()) and same_bits(Ua, Uo)
        a.residual()   # eager work between the replays
    a.close()
    b.close()


@pytest.mark.parametrize("case", ["weno5-visc-rk3", "muscl-rk2-walls", "fp32-weno5-rk3", "2d-weno5-visc-rk3",
                                  "2d-muscl-rk2-walls", "fp32-learned-rk3", "2d-fp32-learned-rk3", "fp64-learned-rk3",
                                  "2d-fp64-learned-rk3", "velin-outlet-last-axis", "2d-inflow-nrbc-last-axis"])
def test_fused_stages_match_separate_passes(oracle, case):
    # the last axis' faces fused with the update (stages 2-3) give the bits of the separate face and update passes,
    # and (FP64) the oracle's
    say("fused stages", case)
    n = (20, 18, 17) if not case.startswith("2d") else (70, 40, 1)
    case = case.replace("2d-", "")
    if case == "weno5-visc-rk3":
        keys = base_keys(*n, reconstruction="weno5", use_viscous="true", mu="0.01", pr="0.71")
        U = smooth_state(*n, 41, mach=0.3)
    elif case == "muscl-rk2-walls":
        keys = base_keys(*n, periodic=False, reconstruction="muscl", time_scheme="SSP Runge-Kutta 2", use_viscous="true",
                         mu="0.01", bc_xmin="wall", bc_xmax="no_slip", bc_ymin="transmissive", bc_ymax="wall",
                         bc_zmin="no_slip", bc_zmax="transmissive", boundary_order="2")
        U = shock_state(*n, 42)
    elif case == "velin-outlet-last-axis":
        # a velocity inlet and a pressure outlet on the last axis (z): their ghost states enter the fused faces
        keys = base_keys(*n, periodic=False, reconstruction="weno5", use_viscous="true", mu="0.01", boundary_order="2",
                         bc_xmin="periodic", bc_xmax="periodic", bc_ymin="wall", bc_ymax="transmissive", bc_zmin="vel_inlet",
                         bc_zmax="outlet", u_in="0.4", v_in="0.1", p_out="0.9")
        U = smooth_state(*n, 45, mach=0.3)
    elif case == "inflow-nrbc-last-axis":
        # a mass-flow inlet and a non-reflecting outlet on the last axis (y in 2-D): the side target and the stored
        # characteristic ghosts enter the fused faces
        keys = base_keys(*n, periodic=False, reconstruction="muscl", boundary_order="2", bc_xmin="periodic", bc_xmax="periodic",
                         bc_ymin="massflow_inlet", bc_ymax="nonreflecting", u_in="0.3", rho_in="0.9", p_out="1.1")
        U = smooth_state(*n, 46, mach=0.3)
    elif case in ("fp32-learned-rk3", "fp64-learned-rk3"):
        # the network's first-stage decisions are used by the fused last-axis faces as by the separate passes; the
        # trained network on a shock state gives both smooth and troubled faces
        keys = base_keys(*n, reconstruction="weno5", flux_precision=case[:4], indicator="learned", indicator_audit="0")
        U = shock_state(*n, 44)
    else:
        keys = base_keys(*n, reconstruction="weno5", flux_precision="fp32", use_viscous="true", mu="0.01")
        U = smooth_state(*n, 43, mach=0.3)
    cfg = fvg.config_from_keys(keys)
    out = []
    for fused in (True, False):
        with fvg.Solver(cfg) as s:
            s.set_fused(fused)
            if case.endswith("learned-rk3"):
                from fenice_amd.train_indicator import DEFAULT_OUT, load
                params, thr, _, _ = load(DEFAULT_OUT)
                s.set_indicator(params, thr)
            s.upload(U)
            t, k, _ = s.advance(0.0, 100.0, max_steps=4)
            out.append((s.download(), t))
    assert out[0][1] == out[1][1] and same_bits(out[0][0], out[1][0])
    if case not in ("fp32-weno5-rk3", "fp32-learned-rk3", "fp64-learned-rk3"):
        Uo = U.copy()
        oracle.run(cfg, Uo, 100.0, max_steps=4)
        assert same_bits(out[0][0], Uo)


@pytest.mark.parametrize("kind", ["tgv3d", "vortex2d"])
def test_monitor_row_matches_the_oracle(oracle, kind):
    # the numbers of a monitor row (SimulationRun.cpp:413-424): volume averages of the state and the RMS norms of its
    # plain residual, against the same sums over the oracle's residual (bit-identical to the device's); the sums
    # differ only in order
    from fenice_amd.workloads import taylor_green, taylor_green_keys
    from cases import VORTEX_HANCOCK_256, vortex_state
    if kind == "tgv3d":
        keys, U = taylor_green_keys(20), taylor_green(20)
    else:
        keys, U = dict(VORTEX_HANCOCK_256, nx="48", ny="48"), vortex_state(48)
    cfg = fvg.config_from_keys(keys)
    with fvg.Solver(cfg) as s:
        s.upload(U)
        s.advance(0.0, 1e9, max_steps=3)
        Ud = s.download()
        row = s.monitor()
    U5 = np.zeros((Ud.shape[0], 5))
    U5[:, :Ud.shape[1]] = Ud
    R, _ = oracle.residual(cfg, U5)
    n = U5.shape[0]
    want = [U5[:, 0].sum() / n, (0.5 * (U5[:, 1] ** 2 + U5[:, 2] ** 2 + U5[:, 4] ** 2) / U5[:, 0]).sum() / n, U5[:, 3].sum() / n,
            np.sqrt((R[:, 0] ** 2).sum() / n), np.sqrt((R[:, 1] ** 2 + R[:, 2] ** 2 + R[:, 4] ** 2).sum() / n),
            np.sqrt((R[:, 3] ** 2).sum() / n)]
    assert np.all(np.abs(row - want) <= 1e-12 * np.abs(want)), (row, want)
    assert row[3] > 0 and row[4] > 0 and row[5] > 0


def quot6_emulated(x):
    """The face kernels' three-operation quotient by 6 (q0 = RN(x a), r = RN(6 q0 - x), RN(q0 - r a) with a = RN(1/6)),
    each rounding done exactly here by rational arithmetic."""
    from fractions import Fraction
    a = 1.0 / 6.0
    q0 = x * a
    r = float(Fraction(q0) * 6 - Fraction(x))
    return float(Fraction(q0) - Fraction(r) * Fraction(a))


def tiny_velocity_state(nx, ny, nz, seed):
    """rho = 1 and u, p of order one; v (and w in 3-D) subnormal multiples of 2^-1074, where quotients by 6 can be
    ties that the three-operation form rounds the wrong way."""
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y = (i.ravel() + 0.5) / nx * 2 * np.pi, (j.ravel() + 0.5) / ny * 2 * np.pi
    n = nx * ny * nz
    tiny = 2.0 ** -1074
    u = 0.3 * np.sin(y)
    v = rng.integers(1, 60, n) * tiny
    w = rng.integers(1, 60, n) * tiny if nz > 1 else np.zeros(n)
    p = 1.0 + 0.1 * np.cos(x)
    U = np.zeros((n, 5))
    U[:, 0] = 1.0
    U[:, 1], U[:, 2], U[:, 4] = u, v, w
    U[:, 3] = p / 0.4 + 0.5 * (u * u + v * v + w * w)
    return U


@pytest.mark.parametrize("dims", [(16, 12, 1), (8, 9, 7)], ids=["2d", "3d"])
def test_weno5_quotients_outside_the_fast_range(oracle, dims):
    # k_cells finds primitives below 2^-900 and the face kernels take the IEEE division for WENO5-Z's quotients by 6
    # (Schemes.cpp:343-345); the state is chosen so that the three-operation quotient would differ on its stencils
    nx, ny, nz = dims
    say(f"weno5 quotients outside the fast range {dims}")
    U = tiny_velocity_state(nx, ny, nz, 5)
    v = U[:, 2].reshape(nz, ny, nx)
    num = [2.0 * np.roll(v, 2, 2) - 7.0 * np.roll(v, 1, 2) + 11.0 * v, -np.roll(v, 1, 2) + 5.0 * v + 2.0 * np.roll(v, -1, 2),
           2.0 * v + 5.0 * np.roll(v, -1, 2) - np.roll(v, -2, 2)]
    differ = sum(float(quot6_emulated(float(x))) != float(x) / 6.0 for a in num for x in a.ravel())
    assert differ > 0   # the fast form would give other bits here
    keys = base_keys(nx, ny, nz, reconstruction="weno5")
    cfg = fvg.config_from_keys(keys)
    two_d = nz == 1
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
        # a state in range afterwards takes the fast form again, with the same bits
        U2 = smooth_state(nx, ny, nz, 6)
        s.upload(U2)
        Rd, Sd = s.residual()
        Ro, So = oracle.residual(cfg, U2)
        assert Sd == So and same_bits(Rd, Ro, two_d)
