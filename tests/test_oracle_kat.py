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


def _quartic_averages(h, n, coef):
    """Cell averages over [i h, (i+1) h] of the quartic sum_k coef[k] x^k, from its antiderivative."""
    x = np.arange(n + 1) * h
    Q = sum(c * x ** (k + 1) / (k + 1) for k, c in enumerate(coef))
    return (Q[1:] - Q[:-1]) / h


@pytest.mark.parametrize("axis", [0, 1])
def test_linear_weights_are_exact_on_quartic_averages(oracle, axis):
    # fvo_set_weno_linear (the smooth-face value of the learned indicator): the upwind-5 face value
    # (2c0 - 13c1 + 47c2 + 27c3 - 3c4)/60 is exact on cell averages of a quartic, from either side. Density and
    # pressure are such averages along one axis and the fluid is at rest, so the face states of an interior face are
    # the point values, the HLLC flux of those (equal) states is (0, p, 0, 0) along the axis, and the momentum residual
    # of an interior cell is -(P(x_{i+1/2}) - P(x_{i-1/2})) / h with P the pressure quartic. WENO5-Z (weights off the
    # linear ones) is not exact there, which shows the check is sharp.
    n, m, L = 32, 4, 2.0
    h = L / n
    prho, pp = (1.0, 0.2, -0.1, 0.03, 0.01), (1.0, 0.3, -0.2, 0.05, 0.02)
    rho_c, p_c = _quartic_averages(h, n, prho), _quartic_averages(h, n, pp)
    shape = (m, n) if axis == 0 else (n, m)   # (ny, nx)
    rho = np.broadcast_to(rho_c if axis == 0 else rho_c[:, None], shape).ravel()
    p = np.broadcast_to(p_c if axis == 0 else p_c[:, None], shape).ravel()
    nx, ny = shape[1], shape[0]
    keys = base_keys(nx, ny, periodic=False, reconstruction="weno5", Lx=str(L if axis == 0 else m * h),
                     Ly=str(m * h if axis == 0 else L), bc_xmin="transmissive", bc_xmax="transmissive", bc_ymin="transmissive",
                     bc_ymax="transmissive", shock_blend="0")
    cfg = config_from_keys(keys)
    U = np.zeros((nx * ny, 5))
    U[:, 0] = rho
    U[:, 3] = p / 0.4
    xf = np.arange(n + 1) * h
    P = sum(c * xf ** k for k, c in enumerate(pp))
    want = -(P[1:] - P[:-1]) / h
    interior = slice(3, n - 3)   # cells whose two faces have full stencils inside the mesh

    def momentum(linear):
        oracle.set_weno_linear(linear)
        try:
            R, _ = oracle.residual(cfg, U)
        finally:
            oracle.set_weno_linear(False)
        R = R.reshape(ny, nx, 5)
        line = R[0, :, 1] if axis == 0 else R[:, 0, 2]
        other = R[..., 2] if axis == 0 else R[..., 1]
        assert np.all(np.abs(other) <= 1e-12 * np.max(np.abs(want)))   # nothing moves across the axis
        return line

    scale = np.max(np.abs(want))
    err_lin = np.max(np.abs(momentum(True)[interior] - want[interior])) / scale
    err_weno = np.max(np.abs(momentum(False)[interior] - want[interior])) / scale
    assert err_lin < 1e-11, err_lin
    assert err_weno > 1e3 * max(err_lin, 1e-15), (err_weno, err_lin)


def test_linear_weights_mirror_symmetric(oracle):
    # the two face values of the linear-weight branch are mirror images of each other: reflecting a jumpy field
    # x -> L - x reflects its residual, with the x momentum's sign flipped (a right value taken from the left stencil
    # would break this)
    nx, ny = 20, 6
    keys = base_keys(nx, ny, periodic=False, reconstruction="weno5", bc_xmin="wall", bc_xmax="wall", bc_ymin="periodic",
                     bc_ymax="periodic")
    cfg = config_from_keys(keys)
    U = shock_state(nx, ny, 1, 19)
    M = U.reshape(ny, nx, 5)[:, ::-1].copy()
    M[..., 1] *= -1.0
    oracle.set_weno_linear(True)
    try:
        R, _ = oracle.residual(cfg, U)
        Rm, _ = oracle.residual(cfg, M.reshape(-1, 5))
    finally:
        oracle.set_weno_linear(False)
    Rm = Rm.reshape(ny, nx, 5)[:, ::-1].copy()
    Rm[..., 1] *= -1.0
    R = R.reshape(ny, nx, 5)
    scale = np.max(np.abs(R), axis=(0, 1))
    assert np.all(np.abs(Rm - R) <= 1e-12 * np.maximum(scale, 1e-30))
