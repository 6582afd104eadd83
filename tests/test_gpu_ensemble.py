"""The ensemble kernels (fvg_batch_*) against the oracle: 1-D instances as the reference's N x 1 runs (its mapping
of a 1-D problem), 2-D instances (isentropic vortices on meshes of uneven shape) as its 2-D runs; every instance's
state, time and step count bit for bit, over schemes, boundaries and sizes down to single cells, independent of the
workgroup size, with per-instance divergence."""
import numpy as np
import pytest

from cases import oracle_instance
from fenice_amd import fvg
from fenice_amd.workloads import ensemble_keys, riemann_ensemble, riemann_states, vortex_ensemble_keys, vortex_rect

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def run_device(keys, n, U, t_end, max_steps, block=None):
    cfg = fvg.config_from_keys(keys)
    with fvg.Ensemble(cfg, n) as e:
        if block:
            e.block = block
        e.upload(U)
        r = e.run(t_end, max_steps)
        return e.download(4), r, e.offsets


def check_against_oracle(oracle, keys, n, U, t_end, max_steps, which=None, block=None):
    Ud, r, off = run_device(keys, n, U, t_end, max_steps, block)
    for i in (range(len(n)) if which is None else which):
        Uo, t, steps, div = oracle_instance(oracle, keys, int(n[i]), U[off[i]:off[i + 1]], t_end, max_steps)
        assert steps == r.steps[i] and div == r.diverged[i], (i, steps, r.steps[i], div, r.diverged[i])
        assert t == r.t[i], (i, t, r.t[i])
        assert same(Ud[off[i]:off[i + 1]], Uo), i
    return r


def test_small_ensemble_matches_oracle(oracle):
    say("ensemble: 40 instances, 6..300 cells, to t = 0.15")
    n, U, _ = riemann_ensemble(40, seed=7, nmin=6, nmax=300)
    r = check_against_oracle(oracle, ensemble_keys(), n, U, 0.15, 2000)
    assert r.diverged.sum() == 0 and (r.t == 0.15).sum() + (r.steps == 2000).sum() == len(n)


SCHEMES = [
    ("muscl-minmod-b1", dict(muscl_limiter="Minmod", boundary_order="1")),
    ("muscl-vanleer-rk3", dict(muscl_limiter="Van Leer", time_scheme="SSP Runge-Kutta 3")),
    ("muscl-koren-euler", dict(muscl_limiter="Koren", time_scheme="Explicit Euler", cfl="0.3")),
    ("weno3", dict(reconstruction="weno3", time_scheme="SSP Runge-Kutta 3")),
    ("weno5-b1", dict(reconstruction="weno5", time_scheme="SSP Runge-Kutta 3", boundary_order="1")),
    ("weno5-b2", dict(reconstruction="weno5", time_scheme="SSP Runge-Kutta 3", boundary_order="2")),
    ("first-order", dict(reconstruction="none", time_scheme="Explicit Euler")),
    ("lowmach-blend", dict(riemann_solver="HLLC (Low-Mach Thornber)", shock_blend="0.2")),
    ("walls", dict(bc_xmin="wall", bc_xmax="no_slip")),
    ("outlet", dict(bc_xmax="outlet", p_out="0.5")),
    ("periodic", dict(bc_xmin="periodic", bc_xmax="periodic", reconstruction="weno5", time_scheme="SSP Runge-Kutta 3")),
    ("stiffened", dict(eos_type="stiffened gas", p_inf="0.5")),
]


@pytest.mark.parametrize("name,over", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_scheme_matrix_matches_oracle(oracle, name, over):
    say("ensemble scheme", name)
    n, _, params = riemann_ensemble(10, seed=11, nmin=1, nmax=90)
    n[:4] = [1, 2, 3, 4]   # below and at the halo depth
    U = riemann_states(n, params)
    check_against_oracle(oracle, dict(ensemble_keys(), **over), n, U, 0.05, 300)


def test_block_size_does_not_change_bits():
    say("ensemble: block sizes")
    n, U, _ = riemann_ensemble(24, seed=3, nmin=16, nmax=1500)
    outs = [run_device(ensemble_keys(), n, U, 0.02, 400, b) for b in (64, 256, 768, 1024)]
    for Ud, r, _ in outs[1:]:
        assert same(Ud, outs[0][0]) and np.array_equal(r.steps, outs[0][1].steps) and same(r.t, outs[0][1].t)


def test_divergence_is_per_instance(oracle):
    say("ensemble: one diverging instance")
    n, U, _ = riemann_ensemble(6, seed=5, nmin=20, nmax=60)
    off = np.concatenate([[0], np.cumsum(n)])
    Ub = U.copy()
    Ub[off[2] + 3] = [1.0, 0.0, 0.0, np.inf]
    Ud, r, _ = run_device(ensemble_keys(), n, Ub, 0.05, 500)
    assert r.diverged.tolist() == [0, 0, 1, 0, 0, 0] and r.steps[2] == 1
    Uc, rc, _ = run_device(ensemble_keys(), n, U, 0.05, 500)
    for i in (0, 1, 3, 4, 5):
        assert same(Ud[off[i]:off[i + 1]], Uc[off[i]:off[i + 1]]) and r.steps[i] == rc.steps[i]
    _, _, steps, div = oracle_instance(oracle, ensemble_keys(), int(n[2]), Ub[off[2]:off[3]], 0.05, 500)
    assert div == 1 and steps == 1


def test_full_size_ensemble_sampled_against_oracle(oracle):
    # the 1024-instance ensemble (128..8192 cells, 4.24 M cells), 30 steps; nine instances spread over the sizes
    say("ensemble: 1024 instances, 30 steps")
    n, U, _ = riemann_ensemble()
    order = np.argsort(n)
    which = [int(order[k]) for k in (0, 1, 128, 300, 512, 700, 900, 1022, 1023)]
    r = check_against_oracle(oracle, ensemble_keys(), n, U, 0.15, 30, which=which)
    assert (r.steps == 30).all() and r.diverged.sum() == 0
    # the default plan splits the largest instances into teams; the sample includes the two largest
    with fvg.Ensemble(fvg.config_from_keys(ensemble_keys()), n) as e:
        e.upload(U)
        e.run(0.15, 1)
        say("  default plan (teams, workgroups in teams, workgroups):", e.plan)
        assert e.plan[0] > 0


# ---- teams: the largest 1-D instances on several workgroups (fvg_batch_set_teams) ------------------------------------
def run_teams(keys, n, U, t_end, max_steps, teams, block=None):
    with fvg.Ensemble(fvg.config_from_keys(keys), n) as e:
        if block:
            e.block = block
        e.set_teams(*teams)
        e.upload(U)
        r = e.run(t_end, max_steps)
        return e.download(4), r, e.offsets, e.plan


@pytest.mark.parametrize("name,over", SCHEMES, ids=[s[0] for s in SCHEMES])
def test_teams_match_oracle(oracle, name, over):
    # every instance of 6 cells or more split over up to 4 workgroups of at least 3 cells (the stencil reach): chunk
    # ends inside the stencils of boundary faces, periodic wrap across the first and last rank, uneven chunks; the
    # same bits as the reference's run of each instance. First-order faces keep one workgroup per instance.
    say("ensemble teams", name)
    keys = dict(ensemble_keys(), **over)
    n, _, params = riemann_ensemble(10, seed=11, nmin=1, nmax=90)
    n[:5] = [1, 2, 5, 6, 7]
    U = riemann_states(n, params)
    Ud, r, off, plan = run_teams(keys, n, U, 0.05, 300, (4, 3, True))
    if keys.get("reconstruction") == "none":
        assert plan[0] == 0
    else:
        assert plan[0] == int((n >= 6).sum()) and plan[1] > plan[0], plan
    for i in range(len(n)):
        Uo, t, steps, div = oracle_instance(oracle, keys, int(n[i]), U[off[i]:off[i + 1]], 0.05, 300)
        assert steps == r.steps[i] and div == r.diverged[i], (i, steps, r.steps[i], div, r.diverged[i])
        assert t == r.t[i], (i, t, r.t[i])
        assert same(Ud[off[i]:off[i + 1]], Uo), i


@pytest.mark.parametrize("block", [64, 256, 768, 1024])
def test_teams_do_not_change_bits(block):
    # teams of up to 16 workgroups of 3 or more cells, and the automatic plan, against one workgroup per instance
    say("ensemble teams: bits, block", block)
    n, U, _ = riemann_ensemble(24, seed=3, nmin=16, nmax=1500)
    for keys in (ensemble_keys(), dict(ensemble_keys(), reconstruction="weno5", time_scheme="SSP Runge-Kutta 3",
                                        bc_xmin="periodic", bc_xmax="periodic")):
        U1, r1, _, plan1 = run_teams(keys, n, U, 0.02, 400, (1, 512, False), block)
        assert plan1[0] == 0
        for teams in ((16, 3, True), (5, 40, True), (8, 64, False)):
            Ut, rt, _, plan = run_teams(keys, n, U, 0.02, 400, teams, block)
            say("  plan", teams, plan)
            if teams[2]:
                assert plan[0] > 0
            assert same(Ut, U1) and np.array_equal(rt.steps, r1.steps) and same(rt.t, r1.t) and same(rt.s_last, r1.s_last)


def test_teamed_divergence_is_reported(oracle):
    # an instance that diverges inside a team: every rank leaves together and the instance reports it
    say("ensemble teams: divergence")
    n, U, _ = riemann_ensemble(6, seed=5, nmin=20, nmax=60)
    off = np.concatenate([[0], np.cumsum(n)])
    Ub = U.copy()
    Ub[off[2] + 3] = [1.0, 0.0, 0.0, np.inf]
    Ud, r, _, plan = run_teams(ensemble_keys(), n, Ub, 0.05, 500, (4, 5, True))
    assert plan[0] == 6
    assert r.diverged.tolist() == [0, 0, 1, 0, 0, 0] and r.steps[2] == 1
    _, _, steps, div = oracle_instance(oracle, ensemble_keys(), int(n[2]), Ub[off[2]:off[3]], 0.05, 500)
    assert div == 1 and steps == 1


@pytest.mark.parametrize("over", [dict(cfl="0.9"), dict(reconstruction="weno5", time_scheme="SSP Runge-Kutta 3", cfl="0.9")],
                         ids=["muscl-rk2", "weno5-rk3"])
def test_teamed_divergence_matches_single_workgroups(over):
    # instances that diverge inside teams, or see extreme values: time, step count, last wave speed, flags and state
    # as with one workgroup per instance
    say("ensemble teams: divergence against single workgroups,", over)
    keys = dict(ensemble_keys(), **over)
    n, U, _ = riemann_ensemble(8, seed=21, nmin=40, nmax=120)
    off = np.concatenate([[0], np.cumsum(n)])
    Ub = U.copy()
    Ub[off[1] + 5] = [1.0, 0.0, 0.0, np.inf]      # the first step's maximum is infinite
    Ub[off[4] + 7] = [1.0, 0.0, 0.0, np.nan]      # a NaN energy (the primitive floors may absorb it)
    Ub[off[6] + 2, 3] = 1e300                     # an extreme energy
    U1, r1, _, plan1 = run_teams(keys, n, Ub, 0.05, 500, (1, 512, False))
    Ut, rt, _, plan = run_teams(keys, n, Ub, 0.05, 500, (4, 10, True))
    assert plan1[0] == 0 and plan[0] == 8
    assert r1.diverged[1] == 1 and r1.diverged.sum() < 8
    assert np.array_equal(r1.diverged, rt.diverged) and np.array_equal(r1.steps, rt.steps)
    assert same(r1.t, rt.t) and same(r1.s_last, rt.s_last) and same(U1, Ut)


def test_team_settings_are_checked():
    with fvg.Ensemble(fvg.config_from_keys(ensemble_keys()), [8, 9]) as e:
        with pytest.raises(fvg.InvalidArgument):
            e.set_teams(0)
        with pytest.raises(fvg.InvalidArgument):
            e.set_teams(17)
        with pytest.raises(fvg.InvalidArgument):
            e.set_teams(4, 2)


def test_tuner_restores_state_and_caches(tmp_path):
    say("ensemble: autotuner")
    n, U, _ = riemann_ensemble(64, seed=9, nmin=64, nmax=2048)
    keys = ensemble_keys()
    cache = str(tmp_path / "tune.txt")
    with fvg.Ensemble(fvg.config_from_keys(keys), n) as e:
        e.upload(U)
        b1, fc1 = e.tune(0.005, 200, cache)
        assert not fc1 and b1 in (64, 128, 256, 512, 768, 1024) and e.block == b1
        assert same(e.download(4), U)
        r = e.run(0.02, 400)
        Ut = e.download(4)
    with fvg.Ensemble(fvg.config_from_keys(keys), n) as e:
        e.upload(U)
        b2, fc2 = e.tune(0.005, 200, cache)
        assert fc2 and b2 == b1
    Ud, r0, _ = run_device(keys, n, U, 0.02, 400, 256)
    assert same(Ud, Ut) and np.array_equal(r.steps, r0.steps)


def test_rejects_unsupported_configurations():
    keys = ensemble_keys()
    with pytest.raises(fvg.Unsupported):
        fvg.Ensemble(fvg.config_from_keys(dict(keys, bc_ymin="wall", bc_ymax="wall")), [8, 9])
    with pytest.raises(fvg.Unsupported):
        fvg.Ensemble(fvg.config_from_keys(dict(keys, use_viscous="true")), [8, 9])
    with pytest.raises(fvg.Unsupported):
        fvg.Ensemble(fvg.config_from_keys(dict(keys, bc_xmin="moving_wall")), [8, 9])
    with fvg.Ensemble(fvg.config_from_keys(keys), [8, 9]) as e:
        with pytest.raises(fvg.LogicError):
            e.run(0.1, 10)
        e.upload(np.ones((17, 3)))
        with pytest.raises(fvg.InvalidArgument):
            e.run(0.1, 0)
        with pytest.raises(fvg.InvalidArgument):
            e.upload(np.ones((16, 3)))


# ---- 2-D instances (fvg_batch_create_2d) ---------------------------------------------------------------------------
def oracle_instance_2d(oracle, keys, nx, ny, U4, t_end, max_steps):
    cfg = fvg.config_from_keys(dict(keys, nx=str(nx), ny=str(ny)))
    U = np.zeros((nx * ny, 5))
    U[:, :4] = U4
    t, steps, diverged = 0.0, 0, 0
    while t < t_end and steps < max_steps:
        dt, _, _, ok = oracle.step(cfg, U, t, t_end, 0.0)
        steps += 1
        if not ok:
            diverged = 1
            break
        t += dt
    return U[:, :4].copy(), t, steps, diverged


SIZES_2D = [(24, 24), (13, 11), (40, 32), (8, 8), (5, 3), (2, 2), (1, 7), (33, 17)]
SCHEMES_2D = [
    ("vortex-muscl-rk2-cfl0.8", dict(cfl="0.8")),   # some instances diverge, as in the oracle
    ("weno5-rk3", dict(reconstruction="weno5", time_scheme="SSP Runge-Kutta 3", cfl="0.5")),
    ("weno3-walls-b2", dict(reconstruction="weno3", time_scheme="SSP Runge-Kutta 3", cfl="0.5", bc_xmin="wall", bc_xmax="no_slip",
                            bc_ymin="transmissive", bc_ymax="outlet", p_out="0.9", boundary_order="2")),
    ("inlet-moving-wall", dict(muscl_limiter="MC", bc_xmin="vel_inlet", bc_xmax="transmissive", bc_ymin="wall",
                               bc_ymax="moving_wall", u_in="0.5", v_in="0.1", boundary_order="2", cfl="0.5")),
    ("first-order-euler", dict(reconstruction="none", time_scheme="Explicit Euler", cfl="0.4")),
    ("lowmach-blend", dict(riemann_solver="HLLC (Low-Mach Thornber)", shock_blend="0.2", muscl_limiter="Koren", cfl="0.5")),
]


@pytest.mark.parametrize("name,over", SCHEMES_2D, ids=[s[0] for s in SCHEMES_2D])
def test_2d_instances_match_oracle(oracle, name, over):
    # isentropic vortices on meshes of uneven shape, down to single rows and columns, each against the oracle's 2-D
    # run of its mesh: state, time and step count bit for bit
    say("ensemble 2-D", name)
    keys = vortex_ensemble_keys(**over)
    nx, ny = np.array([s[0] for s in SIZES_2D]), np.array([s[1] for s in SIZES_2D])
    U = np.concatenate([vortex_rect(a, b) for a, b in SIZES_2D])
    t_end, max_steps = 1.0, 40
    with fvg.Ensemble(fvg.config_from_keys(keys), nx, ny=ny) as e:
        e.upload(U)
        r = e.run(t_end, max_steps)
        Ud = e.download(4)
        off = e.offsets
    for i, (a, b) in enumerate(SIZES_2D):
        Uo, t, steps, div = oracle_instance_2d(oracle, keys, a, b, U[off[i]:off[i + 1]], t_end, max_steps)
        assert steps == r.steps[i] and div == r.diverged[i], (i, steps, r.steps[i], div, r.diverged[i])
        assert t == r.t[i], (i, t, r.t[i])
        assert same(Ud[off[i]:off[i + 1]], Uo), (i, a, b)


@pytest.mark.parametrize("name,over", SCHEMES_2D, ids=[s[0] for s in SCHEMES_2D])
def test_2d_teams_match_oracle(oracle, name, over):
    # every instance of 6 cells and 3 rows or more split into slabs of at least 3 rows over up to 4 workgroups
    # (periodic wrap across the first and last rank, walls, inlets and ghost slabs at the ends, a one-column mesh,
    # uneven slabs): state, time and step count as the oracle's 2-D run of each mesh. First-order faces keep one
    # workgroup per instance.
    say("ensemble 2-D teams", name)
    keys = vortex_ensemble_keys(**over)
    nx, ny = np.array([s[0] for s in SIZES_2D]), np.array([s[1] for s in SIZES_2D])
    U = np.concatenate([vortex_rect(a, b) for a, b in SIZES_2D])
    t_end, max_steps = 1.0, 40
    with fvg.Ensemble(fvg.config_from_keys(keys), nx, ny=ny) as e:
        e.set_teams(4, 3, True)
        e.upload(U)
        r = e.run(t_end, max_steps)
        Ud = e.download(4)
        off = e.offsets
        plan = e.plan
    say("  plan", plan)
    if keys.get("reconstruction") == "none":
        assert plan[0] == 0
    else:
        assert plan[0] == 6 and plan[1] == 19, plan
    for i, (a, b) in enumerate(SIZES_2D):
        Uo, t, steps, div = oracle_instance_2d(oracle, keys, a, b, U[off[i]:off[i + 1]], t_end, max_steps)
        assert steps == r.steps[i] and div == r.diverged[i], (i, steps, r.steps[i], div, r.diverged[i])
        assert t == r.t[i], (i, t, r.t[i])
        assert same(Ud[off[i]:off[i + 1]], Uo), (i, a, b)


@pytest.mark.parametrize("block", [256, 768, 1024])
def test_2d_teams_do_not_change_bits(block):
    # larger vortices on teams of up to 16 workgroups and on the automatic plan, against one workgroup per instance
    say("ensemble 2-D teams: bits, block", block)
    keys = vortex_ensemble_keys(reconstruction="weno5", time_scheme="SSP Runge-Kutta 3", cfl="0.5")
    sizes = [(96, 80), (64, 64), (40, 72), (24, 24)]
    nx, ny = np.array([s[0] for s in sizes]), np.array([s[1] for s in sizes])
    U = np.concatenate([vortex_rect(a, b) for a, b in sizes])
    outs = []
    for teams in ((1, 512, False), (16, 3, True), (8, 256, False)):
        with fvg.Ensemble(fvg.config_from_keys(keys), nx, ny=ny) as e:
            e.block = block
            e.set_teams(*teams)
            e.upload(U)
            r = e.run(0.5, 60)
            say("  plan", teams, e.plan)
            if teams[2]:
                assert e.plan[0] == 4
            outs.append((e.download(4), r))
    for Ud, r in outs[1:]:
        assert same(Ud, outs[0][0]) and np.array_equal(r.steps, outs[0][1].steps) and same(r.t, outs[0][1].t)


def test_2d_block_sizes_and_tuner(oracle):
    # the workgroup size leaves the bits unchanged; the tuner restores the uploaded state
    keys = vortex_ensemble_keys()
    sizes = [(40, 40), (17, 29), (64, 48)]
    nx, ny = np.array([s[0] for s in sizes]), np.array([s[1] for s in sizes])
    U = np.concatenate([vortex_rect(a, b) for a, b in sizes])
    outs = []
    for block in (64, 256, 768, 1024):
        with fvg.Ensemble(fvg.config_from_keys(keys), nx, ny=ny) as e:
            e.block = block
            e.upload(U)
            if block == 1024:
                e.tune(0.05, 100)
                assert same(e.download(4), U)
            e.run(0.5, 30)
            outs.append(e.download(4))
    assert same(outs[0], outs[1]) and same(outs[0], outs[2])


def test_2d_rejects_unsupported_sides():
    with pytest.raises(fvg.Unsupported):
        fvg.Ensemble(fvg.config_from_keys(vortex_ensemble_keys(bc_xmin="massflow_inlet", bc_xmax="transmissive")), [8], ny=[8])
    with pytest.raises(fvg.Unsupported):
        fvg.Ensemble(fvg.config_from_keys(vortex_ensemble_keys(use_viscous="true")), [8], ny=[8])
