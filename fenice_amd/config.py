"""ctypes mirror of `fvg_config` (include/fvg.h) and its enums.

The fields are the reference's run keys for the uniform-grid path (cpp/src/Params.cpp:98-223, typed by
SimParams::from); `Config.from_keys` goes through the library's own parser (fvg_config_set), which accepts the
reference's key names, labels and aliases.
"""
from __future__ import annotations

import ctypes as C

# enum fvg_bc: the order of the reference's BCType (cpp/include/fv/Boundaries.hpp:15)
BC_WALL, BC_NO_SLIP, BC_MOVING_WALL, BC_VEL_INLET, BC_MASSFLOW_INLET, BC_OUTLET, BC_TRANSMISSIVE, BC_PERIODIC, BC_NONREFLECTING = range(9)
LEFT, RIGHT, BOTTOM, TOP, FRONT, BACK = range(6)
HLLC, LOWMACH_HLLC = 0, 1
RECON_NONE, RECON_MUSCL, RECON_WENO3, RECON_WENO5 = range(4)
LIM_MINMOD, LIM_VAN_LEER, LIM_MC, LIM_KOREN = range(4)
TS_AUTO, TS_EULER, TS_SSPRK2, TS_SSPRK3, TS_HANCOCK = -1, 0, 1, 2, 3
EOS_IDEAL, EOS_STIFFENED = 0, 1
FLUX_FP64, FLUX_FP32 = 0, 1
INDICATOR_NONE, INDICATOR_LEARNED = 0, 1
LABEL_WENO5, LABEL_MUSCL, LABEL_BOTH = 0, 1, 2
QUAD_MIDPOINT, QUAD_FOURTH = 0, 1

# status codes
OK, EINVAL, ELOGIC, EDIVERGED, EHIP, ENOMEM, EUNSUPPORTED, EIO = range(8)


class Config(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("eos_type", C.c_int), ("gamma", C.c_double), ("p_inf", C.c_double),
        ("bc", C.c_int * 6),
        ("u_in", C.c_double), ("v_in", C.c_double), ("rho_in", C.c_double), ("p_out", C.c_double),
        ("riemann", C.c_int), ("shock_blend", C.c_double),
        ("reconstruction", C.c_int), ("limiter", C.c_int), ("weno_epsilon", C.c_double), ("boundary_order", C.c_int),
        ("time_scheme", C.c_int), ("cfl", C.c_double),
        ("use_viscous", C.c_int), ("mu", C.c_double), ("pr", C.c_double),
        ("t_end", C.c_double), ("dt_max", C.c_double), ("max_steps", C.c_int),
        ("flux_precision", C.c_int),
        ("indicator", C.c_int), ("indicator_audit", C.c_int), ("indicator_label", C.c_double),
        ("indicator_fn_limit", C.c_double), ("indicator_fn_floor", C.c_double),
        ("flux_quadrature", C.c_int),
        ("nrbc_sigma", C.c_double), ("nrbc_beta", C.c_double),
        ("indicator_label_source", C.c_int),
    ]

    @staticmethod
    def defaults() -> "Config":
        """The reference's baseline defaults (Params.cpp parameter_table) for the keys of the path."""
        c = Config()
        c.nx, c.ny, c.nz = 0, 0, 1
        c.Lx, c.Ly, c.Lz = 1.0, 1.0, 1.0
        c.eos_type, c.gamma, c.p_inf = EOS_IDEAL, 1.4, 0.0
        # the baseline sides of BoundaryManager (Boundaries.hpp:128-129): transmissive x, slip walls y, periodic z
        for s, b in enumerate([BC_TRANSMISSIVE, BC_TRANSMISSIVE, BC_WALL, BC_WALL, BC_PERIODIC, BC_PERIODIC]):
            c.bc[s] = b
        c.u_in, c.v_in, c.rho_in, c.p_out = 0.0, 0.0, 1.0, 1.0
        c.riemann, c.shock_blend = HLLC, 0.2
        c.reconstruction, c.limiter, c.weno_epsilon, c.boundary_order = RECON_NONE, LIM_VAN_LEER, 1e-6, 2
        c.time_scheme, c.cfl = TS_AUTO, 0.5
        c.use_viscous, c.mu, c.pr = 0, 1e-3, 0.72
        c.t_end, c.dt_max, c.max_steps = 0.005, 0.0, 0
        c.flux_precision = FLUX_FP64
        c.indicator, c.indicator_audit, c.indicator_label, c.indicator_fn_limit = INDICATOR_NONE, 50, 0.05, 0.01
        c.indicator_fn_floor = 1e-4
        c.flux_quadrature = QUAD_MIDPOINT
        c.nrbc_sigma, c.nrbc_beta = 0.28, -1.0
        c.indicator_label_source = LABEL_WENO5
        return c

    @staticmethod
    def make(**kw) -> "Config":
        """Defaults with typed fields overridden; `bc` takes a 6-sequence, `periodic=True` makes every side periodic."""
        c = Config.defaults()
        if kw.pop("periodic", False):
            for s in range(6):
                c.bc[s] = BC_PERIODIC
        bc = kw.pop("bc", None)
        if bc is not None:
            for s, b in enumerate(bc):
                c.bc[s] = b
        for k, v in kw.items():
            if not hasattr(c, k):
                raise KeyError(k)
            setattr(c, k, v)
        if c.time_scheme == TS_AUTO:
            c.time_scheme = (TS_EULER if c.reconstruction != RECON_MUSCL
                             else (TS_SSPRK2 if c.use_viscous else TS_HANCOCK))
        return c

    @property
    def ncells(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def is3d(self) -> bool:
        return self.nz > 1

    def copy(self) -> "Config":
        c = Config()
        C.memmove(C.byref(c), C.byref(self), C.sizeof(Config))
        return c
