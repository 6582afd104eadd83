"""The Kelvin-Helmholtz shear layer (periodic [0, 1]^2, WENO5-Z + HLLC + SSP-RK3, CFL 0.5) up to 2048^2 cells:
steps bit-identical to the oracle, and at the full size the properties the periodic problem fixes: equivariance
under a periodic shift of the mesh (bit for bit) and conservation of mass and energy to round-off."""
import numpy as np
import pytest

from fenice_amd import fvg
from fenice_amd.workloads import kelvin_helmholtz, kh_keys

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def device_run(n, U, steps):
    with fvg.Solver(fvg.config_from_keys(kh_keys(n))) as s:
        s.upload(U)
        tot0 = s.conserved_totals()
        t, k, _ = s.advance(0.0, 100.0, max_steps=steps)
        return s.download(4), t, tot0, s.conserved_totals()


@pytest.mark.parametrize("n,steps", [(128, 12), (2048, 2)])
def test_kh_matches_oracle(oracle, n, steps):
    say(f"kh {n}^2, {steps} steps")
    U = kelvin_helmholtz(n)
    Ud, t, _, _ = device_run(n, U, steps)
    Uo = np.zeros((n * n, 5))
    Uo[:, :4] = U
    to, k = oracle.run(fvg.config_from_keys(kh_keys(n)), Uo, 100.0, max_steps=steps)
    assert k == steps and t == to
    assert same(Ud, Uo[:, :4])


def test_kh_2048_shift_equivariance_and_conservation():
    n = 2048
    say("kh 2048^2 shift / conservation")
    U = kelvin_helmholtz(n)
    sx, sy = 37, 611
    Us = np.roll(U.reshape(n, n, 4), (sy, sx), axis=(0, 1)).reshape(-1, 4)
    Ua, ta, tot0, tot1 = device_run(n, U, 4)
    Ub, tb, _, _ = device_run(n, Us, 4)
    assert ta == tb
    assert same(np.roll(Ua.reshape(n, n, 4), (sy, sx), axis=(0, 1)).reshape(-1, 4), Ub)
    assert abs(tot1[0] - tot0[0]) <= 1e-13 * abs(tot0[0])
    assert abs(tot1[2] - tot0[2]) <= 1e-13 * abs(tot0[2])
