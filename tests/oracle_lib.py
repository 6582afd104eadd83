"""Loaders of the oracle (oracle/liboracle.so) and of the reference built from its sources
(oracle/_ref/libfenice_ref.so) for the tests. Test infrastructure: the checker, never the thing measured."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from fenice_amd.config import Config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libfenice_ref.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def _dp(a):
    return a.ctypes.data_as(_D)


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        self.lib = L = C.CDLL(path)
        P = C.POINTER(Config)
        L.fvo_dt_length.argtypes = [P]
        L.fvo_dt_length.restype = C.c_double
        L.fvo_face_flux_batch.argtypes = [P, _D, _D, _I, C.c_size_t, _D, _D]
        L.fvo_residual.argtypes = [P, _D, _D, C.c_double, _D, C.POINTER(C.c_longlong)]
        L.fvo_residual.restype = C.c_int
        L.fvo_step.argtypes = [P, _D, C.c_double, C.c_double, C.c_double, _D, _D, _D]
        L.fvo_step.restype = C.c_int
        L.fvo_volume_averages.argtypes = [P, _D, _D]
        L.fvo_set_threads.argtypes = [C.c_int]
        L.fvo_get_threads.restype = C.c_int

    def set_weno_linear(self, on: bool):
        self.lib.fvo_set_weno_linear(int(on))

    def set_threads(self, n: int):
        self.lib.fvo_set_threads(n)

    def dt_length(self, cfg: Config) -> float:
        return self.lib.fvo_dt_length(C.byref(cfg))

    def face_flux(self, cfg, UL, UR, axis):
        UL = np.ascontiguousarray(UL, np.float64)
        UR = np.ascontiguousarray(UR, np.float64)
        ax = np.ascontiguousarray(axis, np.int32)
        n = UL.shape[0]
        F = np.empty((n, 5))
        S = np.empty(n)
        self.lib.fvo_face_flux_batch(C.byref(cfg), _dp(UL), _dp(UR), ax.ctypes.data_as(_I), n, _dp(F), _dp(S))
        return F, S

    @staticmethod
    def _check(cfg, U):
        if U.ndim != 2 or U.shape[1] != 5 or U.shape[0] != cfg.nx * cfg.ny * cfg.nz:
            raise ValueError(f"the oracle takes a ({cfg.nx * cfg.ny * cfg.nz}, 5) state; got {U.shape}")

    def residual(self, cfg, U, half_dt=0.0):
        U = np.ascontiguousarray(U, np.float64)
        self._check(cfg, U)
        R = np.empty_like(U)
        S = C.c_double()
        rev = C.c_longlong()
        rc = self.lib.fvo_residual(C.byref(cfg), _dp(U), _dp(R), half_dt, C.byref(S), C.byref(rev))
        if rc != 0:
            raise RuntimeError(f"fvo_residual failed ({rc})")
        return R, S.value

    def fourth_stats(self, cfg, U):
        """Branch counts of the fourth-order residual of U: dict of faces reconstructed, fallbacks (admissibility),
        good, pt, full, pt_only (pt without full), good_only (good without pt)."""
        U = np.ascontiguousarray(U, np.float64)
        out = (C.c_longlong * 7)()
        rc = self.lib.fvo_fourth_stats(C.byref(cfg), _dp(U), out)
        if rc != 0:
            raise RuntimeError(f"fvo_fourth_stats failed ({rc})")
        return dict(zip(("reconstructed", "fallbacks", "good", "pt", "full", "pt_only", "good_only"), list(out)))

    def step(self, cfg, U, t, t_max, dt_max=0.0, lagged=0.0):
        """One step in place on U; returns (dt, s_max, lagged_speed, ok)."""
        self._check(cfg, U)
        if not (U.flags.c_contiguous and U.dtype == np.float64):
            raise ValueError("the oracle steps a C-contiguous float64 state in place")
        lag = C.c_double(lagged)
        dt = C.c_double()
        S = C.c_double()
        rc = self.lib.fvo_step(C.byref(cfg), _dp(U), t, t_max, dt_max, C.byref(lag), C.byref(dt), C.byref(S))
        if rc not in (0, -1):
            raise RuntimeError(f"fvo_step failed ({rc})")
        return dt.value, S.value, lag.value, rc == 0

    def run(self, cfg, U, t_end, max_steps=0):
        """Steps while t < t_end (Simulation::run_loop); returns (t, steps)."""
        t, steps, lag = 0.0, 0, 0.0
        while t < t_end:
            dt, S, lag, ok = self.step(cfg, U, t, t_end, 0.0, lag)
            steps += 1
            if not ok:
                raise RuntimeError(f"diverged at step {steps}")
            t += dt
            if max_steps > 0 and steps >= max_steps:
                break
        return t, steps

    def volume_averages(self, cfg, U):
        out = np.empty(3)
        self.lib.fvo_volume_averages(C.byref(cfg), _dp(np.ascontiguousarray(U, np.float64)), _dp(out))
        return out


class Reference:
    """The reference's own Simulation (built from its sources) for a test case and key overrides."""

    def __init__(self, path: str = REF_SO):
        os.environ.setdefault("FENICE_NO_SIMD", "1")
        self.lib = L = C.CDLL(path)
        L.ref_last_error.restype = C.c_char_p
        L.ref_create.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p)]
        L.ref_create.restype = C.c_void_p
        L.ref_destroy.argtypes = [C.c_void_p]
        L.ref_num_cells.argtypes = [C.c_void_p]
        L.ref_num_cells.restype = C.c_longlong
        L.ref_dt_length.argtypes = [C.c_void_p]
        L.ref_dt_length.restype = C.c_double
        L.ref_initial_field.argtypes = [C.c_void_p, _D]
        L.ref_set_state.argtypes = [C.c_void_p, _D]
        L.ref_get_state.argtypes = [C.c_void_p, _D]
        L.ref_residual.argtypes = [C.c_void_p, _D, C.c_double, _D, _D]
        L.ref_step.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, _D, _D]
        L.ref_run.argtypes = [C.c_void_p, _D, C.c_double, C.c_int]
        L.ref_get_lagged_speed.argtypes = [C.c_void_p]
        L.ref_get_lagged_speed.restype = C.c_double
        L.ref_set_lagged_speed.argtypes = [C.c_void_p, C.c_double]
        L.ref_face_flux.argtypes = [C.c_double, C.c_int, C.c_double, C.c_int, C.c_double, _D, _D, _I, C.c_longlong, _D, _D]
        L.ref_checkpoint_path.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.ref_save_checkpoint.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.ref_load_checkpoint.argtypes = [C.c_void_p, _D]
        L.ref_volume_averages.argtypes = [C.c_void_p, _D, _D]

    def create(self, case: str, **keys) -> "RefRun":
        ks = [k.encode() for k in keys]
        vs = [str(v).encode() for v in keys.values()]
        h = self.lib.ref_create(case.encode(), len(ks), (C.c_char_p * len(ks))(*ks), (C.c_char_p * len(vs))(*vs))
        if not h:
            raise RuntimeError(self.lib.ref_last_error().decode())
        return RefRun(self, h)

    def face_flux(self, UL, UR, axis, gamma=1.4, stiffened=False, p_inf=0.0, lowmach=False, blend=0.0):
        UL = np.ascontiguousarray(UL, np.float64)
        UR = np.ascontiguousarray(UR, np.float64)
        ax = np.ascontiguousarray(axis, np.int32)
        n = UL.shape[0]
        F = np.empty((n, 5))
        S = np.empty(n)
        self.lib.ref_face_flux(gamma, int(stiffened), p_inf, int(lowmach), blend, _dp(UL), _dp(UR), ax.ctypes.data_as(_I), n,
                               _dp(F), _dp(S))
        return F, S


class RefRun:
    def __init__(self, ref: Reference, h):
        self.ref, self.h, self.lib = ref, h, ref.lib
        self.n = self.lib.ref_num_cells(h)

    def __del__(self):
        try:
            self.lib.ref_destroy(self.h)
        except Exception:
            pass

    @property
    def dt_length(self):
        return self.lib.ref_dt_length(self.h)

    def initial_field(self):
        U = np.empty((self.n, 5))
        if self.lib.ref_initial_field(self.h, _dp(U)) != 0:
            raise RuntimeError(self.lib.ref_last_error().decode())
        return U

    def set_state(self, U):
        self.lib.ref_set_state(self.h, _dp(np.ascontiguousarray(U, np.float64)))

    def get_state(self):
        U = np.empty((self.n, 5))
        self.lib.ref_get_state(self.h, _dp(U))
        return U

    def residual(self, U, half_dt=0.0):
        U = np.ascontiguousarray(U, np.float64)
        R = np.empty_like(U)
        S = C.c_double()
        if self.lib.ref_residual(self.h, _dp(U), half_dt, _dp(R), C.byref(S)) != 0:
            raise RuntimeError(self.lib.ref_last_error().decode())
        return R, S.value

    def step(self, t, t_max, dt_max=0.0):
        dt, S = C.c_double(), C.c_double()
        if self.lib.ref_step(self.h, t, t_max, dt_max, C.byref(dt), C.byref(S)) != 0:
            raise RuntimeError(self.lib.ref_last_error().decode())
        return dt.value, S.value

    def volume_averages(self, U):
        """Simulation::volume_averages of U (the reference's own sequential sums)."""
        out = np.empty(3)
        if self.lib.ref_volume_averages(self.h, _dp(np.ascontiguousarray(U, np.float64)), _dp(out)) != 0:
            raise RuntimeError(self.lib.ref_last_error().decode())
        return out

    def run(self, t_end, max_steps=0):
        t = C.c_double(0.0)
        steps = self.lib.ref_run(self.h, C.byref(t), t_end, max_steps)
        return t.value, steps

    @property
    def checkpoint_path(self) -> str:
        buf = C.create_string_buffer(4096)
        if self.lib.ref_checkpoint_path(self.h, buf, 4096) < 0:
            raise RuntimeError("checkpoint path too long")
        return buf.value.decode()

    def save_checkpoint(self, t, step):
        """Simulation::save_checkpoint (the reference's own writer) into checkpoint_path."""
        if self.lib.ref_save_checkpoint(self.h, t, step) != 0:
            raise RuntimeError(self.lib.ref_last_error().decode())

    def load_checkpoint(self):
        """Simulation::load_checkpoint (the reference's own reader) from checkpoint_path: (t, step)."""
        t = C.c_double()
        step = self.lib.ref_load_checkpoint(self.h, C.byref(t))
        if step < 0:
            raise RuntimeError(self.lib.ref_last_error().decode())
        return t.value, step

    @property
    def lagged_speed(self):
        return self.lib.ref_get_lagged_speed(self.h)

    @lagged_speed.setter
    def lagged_speed(self, s):
        self.lib.ref_set_lagged_speed(self.h, s)
