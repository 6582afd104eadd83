"""The oracle against the reference's stored regression states (cpp/tests/reference/*.npz, copied into
tests/golden with the initial field the reference makes): same step count, max relative difference below the
reference's own tolerance 1e-8 (cpp/tests/regression.py:295-305)."""
import numpy as np
import pytest

from cases import golden_names, load_golden
from fenice_amd.fvg import config_from_keys


def rel_diff(U, U_ref):
    scale = np.maximum(np.max(np.abs(U_ref), axis=0), 1e-30)
    return float(np.max(np.abs(U - U_ref) / scale))


@pytest.mark.parametrize("name", golden_names())
def test_oracle_reproduces_reference_state(oracle, name):
    g = load_golden(name)
    cfg = config_from_keys(g["keys"])
    U = g["U0"].copy()
    t, steps = oracle.run(cfg, U, float(g["keys"]["t_end"]))
    assert steps == g["step"]
    assert rel_diff(U[:, : g["U_ref"].shape[1]], g["U_ref"]) < 1e-8
