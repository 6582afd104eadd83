"""The C ABI without a device: the library loads, exports every entry point include/fvg.h declares, and parses
the reference's keys, labels and aliases with the reference's error behaviour."""
import ctypes as C
import os
import re

import pytest

from fenice_amd import fvg
from fenice_amd.config import (BC_MOVING_WALL, BC_NO_SLIP, BC_PERIODIC, BC_TRANSMISSIVE, BC_WALL, HLLC, LIM_KOREN, LIM_MC,
                               LIM_MINMOD, LIM_VAN_LEER, LOWMACH_HLLC, RECON_MUSCL, RECON_NONE, RECON_WENO3, RECON_WENO5, TS_EULER,
                               TS_HANCOCK, TS_SSPRK2, TS_SSPRK3, Config)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "fvg.h")).read()
    return sorted(set(re.findall(r"^(?:int|void|double|const char\s*\*)\s*(fvg_[a-z_0-9]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_entry_point():
    L = fvg.lib()
    names = declared_entry_points()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert set(fvg.EXPORTS) <= set(names)


def test_config_struct_size_matches_header():
    # the ctypes mirror and the C struct agree: defaults written by C read back field by field
    c = Config()
    fvg.lib().fvg_config_defaults(C.byref(c))
    d = Config.defaults()
    for f, _ in Config._fields_:
        if f == "bc":
            assert list(c.bc) == list(d.bc)
        elif f != "time_scheme":
            assert getattr(c, f) == getattr(d, f), f


@pytest.mark.parametrize("label,expect", [("HLLC", HLLC), ("hllc", HLLC), ("HLLC (Low-Mach Thornber)", LOWMACH_HLLC),
                                          ("lowmach", LOWMACH_HLLC), ("Low-Mach", LOWMACH_HLLC)])
def test_riemann_labels(label, expect):
    assert fvg.config_from_keys(dict(nx=4, ny=4, riemann_solver=label)).riemann == expect


def test_labels_and_auto_scheme():
    c = fvg.config_from_keys(dict(nx=4, ny=4, reconstruction="muscl", muscl_limiter="Van Leer"))
    assert c.reconstruction == RECON_MUSCL and c.limiter == LIM_VAN_LEER and c.time_scheme == TS_HANCOCK
    c = fvg.config_from_keys(dict(nx=4, ny=4, reconstruction="MUSCL", use_viscous="true"))
    assert c.time_scheme == TS_SSPRK2
    c = fvg.config_from_keys(dict(nx=4, ny=4, reconstruction="weno5"))
    assert c.reconstruction == RECON_WENO5 and c.time_scheme == TS_EULER
    c = fvg.config_from_keys(dict(nx=4, ny=4))
    assert c.reconstruction == RECON_NONE and c.time_scheme == TS_EULER
    for lab, lim in [("Minmod", LIM_MINMOD), ("vanleer", LIM_VAN_LEER), ("MC", LIM_MC), ("koren", LIM_KOREN)]:
        assert fvg.config_from_keys(dict(nx=4, ny=4, muscl_limiter=lab)).limiter == lim
    for lab, ts in [("SSP Runge-Kutta 3", TS_SSPRK3), ("rk2", TS_SSPRK2), ("Explicit Euler", TS_EULER), ("hancock", TS_HANCOCK)]:
        assert fvg.config_from_keys(dict(nx=4, ny=4, reconstruction="muscl", time_scheme=lab)).time_scheme == ts


def test_scheme_combinations_checked_like_the_reference():
    # SimParams::from (Params.cpp:429-460): the same rejections and messages, and the order-1 normalisation
    with pytest.raises(fvg.InvalidArgument, match="weno5 is not used with SSP Runge-Kutta 2"):
        fvg.config_from_keys(dict(nx=4, ny=4, reconstruction="weno5", time_scheme="SSP Runge-Kutta 2"))
    with pytest.raises(fvg.InvalidArgument, match="weno3 is not used with SSP Runge-Kutta 2"):
        fvg.config_from_keys(dict(nx=4, ny=4, reconstruction="weno3", time_scheme="rk2"))
    with pytest.raises(fvg.InvalidArgument, match="MUSCL-Hancock needs reconstruction = muscl"):
        fvg.config_from_keys(dict(nx=4, ny=4, reconstruction="weno5", time_scheme="MUSCL-Hancock"))
    with pytest.raises(fvg.InvalidArgument, match="MUSCL-Hancock needs reconstruction = muscl"):
        fvg.config_from_keys(dict(nx=4, ny=4, time_scheme="hancock"))
    with pytest.raises(fvg.InvalidArgument, match="weno_epsilon must be >= 0"):
        fvg.config_from_keys(dict(nx=4, ny=4, weno_epsilon="-1"))
    assert fvg.config_from_keys(dict(nx=4, ny=4, boundary_order="2")).boundary_order == 1
    assert fvg.config_from_keys(dict(nx=4, ny=4, reconstruction="muscl", boundary_order="2")).boundary_order == 2
    assert fvg.config_from_keys(dict(nx=4, ny=4, reconstruction="weno3")).reconstruction == RECON_WENO3
    c = fvg.config_from_keys(dict(nx=4, ny=4, bc_xmin="Inviscid Wall (Slip)", bc_xmax="no_slip", bc_ymin="Moving Wall",
                                  bc_ymax="transmissive"))
    assert list(c.bc)[:4] == [BC_WALL, BC_NO_SLIP, BC_MOVING_WALL, BC_TRANSMISSIVE]
    c = fvg.config_from_keys(dict(nx=4, ny=4, nz=4, bc_zmin="periodic", bc_zmax="periodic"))
    assert c.bc[4] == BC_PERIODIC and c.is3d


@pytest.mark.parametrize("keys,msg", [
    (dict(nx=4, ny=4, riemann_solver="roe"), "Unknown Riemann solver"),
    (dict(nx=4, ny=4, riemann_solver="HLLC (Preconditioned)"), "removed"),
    (dict(nx=4, ny=4, muscl_limiter="superbee"), "Unknown MUSCL limiter"),
    (dict(nx=4, ny=4, reconstruction="weno7"), "Unknown reconstruction"),
    (dict(nx=4, ny=4, time_scheme="rk4"), "Unknown time scheme"),
    (dict(nx=4, ny=4, bc_xmin="sticky"), "Unknown boundary condition"),
    (dict(nx=4, ny=4, cfl="fast"), "not a number"),
    (dict(nx=4, ny=4, no_such_key=1), "Unknown key"),
    (dict(nx=0, ny=4), "positive"),
    (dict(nx=4, ny=4, bc_xmin="periodic", bc_xmax="wall"), "pairs"),
    # SimParams::from (Params.cpp:476-478, 491-496): the two EOS labels only, and the run keys' ranges
    (dict(nx=4, ny=4, eos_type="ideal", p_inf="1"), "Unknown eos_type"),
    (dict(nx=4, ny=4, eos_type="tait"), "Unknown eos_type"),
    (dict(nx=4, ny=4, cfl="0"), "cfl must be positive"),
    (dict(nx=4, ny=4, cfl="-0.5"), "cfl must be positive"),
    (dict(nx=4, ny=4, t_end="0"), "t_end must be positive"),
    (dict(nx=4, ny=4, dt_max="-1e-3"), "dt_max must be finite and >= 0"),
    (dict(nx=4, ny=4, dt_max="inf"), "dt_max must be finite and >= 0"),
    (dict(nx=4, ny=4, gamma="1.0"), "gamma must be greater than 1"),
    # the learned indicator (no reference key): WENO5-Z, a Runge-Kutta scheme, the midpoint quadrature
    (dict(nx=4, ny=4, indicator="learned", reconstruction="weno3"), "indicator = learned needs reconstruction = weno5"),
    (dict(nx=4, ny=4, indicator="learned", reconstruction="weno5", flux_quadrature="fourth"),
     "indicator = learned needs flux_quadrature = midpoint"),
])
def test_invalid_keys_raise_value_error(keys, msg):
    with pytest.raises(ValueError, match=msg):
        fvg.config_from_keys(keys)


def test_reference_case_keys_outside_the_path():
    # a whole case file's keys (Params.cpp parameter_table): run control and inert feature parameters are accepted,
    # a switched-on feature outside the path is refused as unsupported
    keys = dict(test_case="Sod Shock Tube", nx=40, ny=4, bc_xmin="transmissive", bc_xmax="transmissive", bc_ymin="wall",
                bc_ymax="wall", save_interval=10, write_interval_time="0.01", monitor_interval=5, resume="false", dump_state="",
                vtk_format="binary", n_cpus=4, use_bodyforce="false", gx="0", gy="-9.81", gz="0", source_rho="0", source_mx="0.0",
                use_ibm="false", ibm_shape="Cylinder", ibm_radius="0.1", use_amr="no", amr_max_level=2, use_mpi="off",
                positivity_fallback="false", reynolds="0", u_in_table="", p_out_table="")
    c = fvg.config_from_keys(keys)
    assert c.nx == 40 and c.bc[0] == BC_TRANSMISSIVE
    for k, v in [("use_ibm", "true"), ("use_amr", "1"), ("use_mpi", "yes"), ("use_bodyforce", "on"),
                 ("positivity_fallback", "true"), ("source_rho", "0.1"), ("source_E", "sin(x)"), ("reynolds", "100"),
                 ("u_in_table", "inflow.csv")]:
        with pytest.raises(fvg.Unsupported, match="outside the uniform-grid device path"):
            fvg.config_from_keys(dict(nx=8, ny=8, **{k: v}))
    with pytest.raises(ValueError, match="not a flag"):
        fvg.config_from_keys(dict(nx=8, ny=8, use_ibm="maybe"))


def test_label_sources():
    from fenice_amd.config import LABEL_BOTH, LABEL_MUSCL, LABEL_WENO5
    for v, want in (("weno5", LABEL_WENO5), ("MUSCL", LABEL_MUSCL), ("both", LABEL_BOTH)):
        assert fvg.config_from_keys(dict(nx=4, ny=4, indicator_label_source=v)).indicator_label_source == want
    assert fvg.config_from_keys(dict(nx=4, ny=4)).indicator_label_source == LABEL_WENO5
    with pytest.raises(ValueError, match="Unknown indicator_label_source"):
        fvg.config_from_keys(dict(nx=4, ny=4, indicator_label_source="tvd"))


def test_learned_indicator_in_both_flux_precisions():
    for prec in ("fp64", "fp32"):
        c = fvg.config_from_keys(dict(nx=4, ny=4, indicator="learned", reconstruction="weno5", flux_precision=prec))
        assert c.indicator == 1 and c.flux_precision == (prec == "fp32")


def test_eos_labels():
    from fenice_amd.config import EOS_STIFFENED
    assert fvg.config_from_keys(dict(nx=4, ny=4, eos_type="Stiffened Gas", p_inf="0.5")).eos_type == EOS_STIFFENED
    assert fvg.config_from_keys(dict(nx=4, ny=4, eos_type="IDEAL GAS")).eos_type == 0


def test_dt_length():
    c = fvg.config_from_keys(dict(nx=8, ny=4, Lx=1.0, Ly=1.0))
    assert fvg.lib().fvg_dt_length(C.byref(c)) == 0.125
    c = fvg.config_from_keys(dict(nx=4, ny=4, nz=4, Lx=1.0, Ly=1.0, Lz=1.0))
    assert fvg.lib().fvg_dt_length(C.byref(c)) == 2.0 / (4.0 + 4.0 + 4.0)
