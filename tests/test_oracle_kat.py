"""The oracle pinned without the reference running: the reference's Riemann fingerprints and invariants the
mathematics fixes (free stream, translation, conservation, the time scheme's order)."""
import numpy as np
import pytest

from cases import base_keys, smooth_state, shock_state
from fenice_amd.config import Config, EOS_STIFFENED, HLLC, LOWMACH_HLLC
from fenice_amd.fvg import config_from_keys
from riemann_pool import HASHES, fingerprint, pool


@pytest.mark.parametrize("kind", ["ideal", "stiffened"])
@pytest.mark.parametrize("solver", ["hllc", "lowmach"])
def test_riemann_fingerprint(oracle, kind, solver):
    UL, UR, ax, gamma, p_inf, stiff = pool(kind)
    cfg = Config.make(gamma=gamma, p_inf=p_inf, eos_type=EOS_STIFFENED if stiff else 0, shock_blend=0.0,
                      riemann=LOWMACH_HLLC if solver == "lowmach" else HLLC)
    F, S = oracle.face_flux(cfg, UL, UR, ax)
    assert fingerprint(F, S) == HASHES[(kind, solver)]


def test_riemann_matches_reference(oracle, ref):
    UL, UR, ax, *_ = pool("ideal")
    for blend in (0.0, 0.2):
        for lm in (0, 1):
            cfg = Config.make(shock_blend=blend, riemann=lm)
            F, S = oracle.face_flux(cfg, UL, UR, ax)
            Fr, Sr = ref.face_flux(UL, UR, ax, lowmach=bool(lm), blend=blend)
            assert np.array_equal(F.view(np.uint64), Fr.view(np.uint64))
            assert np.array_equal(S.view(np.uint64), Sr.view(np.uint64))


@pytest.mark.parametrize("rc", ["muscl", "weno3", "weno5", "none"])
def test_free_stream_is_preserved(oracle, rc):
    # a uniform state has zero residual, exactly (every face flux cancels against its neighbour)
    keys = base_keys(6, 5, 4, reconstruction=rc, use_viscous="true", mu="0.01")
    cfg = config_from_keys(keys)
    U = np.tile(np.array([1.3, 0.4, -0.2, 3.1, 0.25]), (cfg.ncells, 1))
    R, S = oracle.residual(cfg, U)
    assert np.all(R == 0.0)
    assert S > 0


@pytest.mark.parametrize("rc,visc", [("muscl", "false"), ("weno5", "true"), ("weno3", "false")])
def test_periodic_translation(oracle, rc, visc):
    # on a periodic mesh, shifting the field shifts the residual bit for bit (test_periodic.cpp:593-617)
    nx, ny, nz = 8, 7, 6
    cfg = config_from_keys(base_keys(nx, ny, nz, reconstruction=rc, use_viscous=visc, mu="0.02"))
    U = smooth_state(nx, ny, nz, 3)
    R, S = oracle.residual(cfg, U)
    sh = (1, 2, 3)
    roll = lambda A: np.roll(A.reshape(nz, ny, nx, 5), sh, axis=(0, 1, 2)).reshape(-1, 5)
    R2, S2 = oracle.residual(cfg, roll(U))
    assert np.array_equal(R2, roll(R)) and S2 == S


def test_periodic_conservation(oracle):
    # sum of the residual over a periodic mesh vanishes to round-off (the divergence telescopes)
    nx, ny, nz = 10, 9, 8
    cfg = config_from_keys(base_keys(nx, ny, nz, reconstruction="weno5", use_viscous="true", mu="0.01"))
    U = shock_state(nx, ny, nz, 5)
    R, _ = oracle.residual(cfg, U)
    scale = np.abs(R).max(axis=0)
    assert np.all(np.abs(R.sum(axis=0)) <= 1e-12 * scale * R.shape[0])


def test_rk_order_on_smooth_flow(oracle):
    # SSP-RK3 is third order in time: halving dt (through dt_max) divides the error against a fine run by ~8
    nx = ny = 16
    U0 = smooth_state(nx, ny, 1, 9, noise=0.0)
    cfg = config_from_keys(base_keys(nx, ny, reconstruction="weno5", cfl=0.5))
    T = 0.02

    def run(dtm):
        U = U0.copy()
        t = 0.0
        while t < T:
            dt, _, _, ok = oracle.step(cfg, U, t, T, dtm)
            assert ok
            t += dt
        return U

    fine = run(T / 64)
    e1 = np.abs(run(T / 4) - fine).max()
    e2 = np.abs(run(T / 8) - fine).max()
    assert 5.0 < e1 / e2 < 12.0
