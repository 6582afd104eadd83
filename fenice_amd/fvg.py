"""Python binding of the device library (libfvg.so) through its C ABI (include/fvg.h).

`Solver` mirrors the reference's step interface for the uniform path: `residual()` is
Simulation::residual (cpp/src/Simulation.cpp:1351) returning (R, S_max), `step()` is TimeScheme::step
(cpp/include/fv/TimeScheme.hpp:48) returning StepResult(dt, s_max), and failures raise the exception kinds
the reference raises (ValueError for std::invalid_argument, LogicError for std::logic_error, Diverged for
the runtime_error of check_progress). There is no CPU fallback: without the built library this module raises
on import of the library.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from .config import Config, EDIVERGED, EHIP, EINVAL, EIO, ELOGIC, ENOMEM, EUNSUPPORTED, OK

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfvg.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)

EXPORTS = [
    "fvg_config_defaults", "fvg_config_set", "fvg_config_finalize", "fvg_dt_length", "fvg_create", "fvg_destroy",
    "fvg_last_error", "fvg_upload", "fvg_download", "fvg_residual", "fvg_step", "fvg_advance", "fvg_get_lagged_speed",
    "fvg_set_lagged_speed", "fvg_conserved_totals", "fvg_kernel_times", "fvg_set_timing", "fvg_set_graphs", "fvg_set_fused", "fvg_riemann_batch",
    "fvg_batch_create", "fvg_batch_destroy", "fvg_batch_last_error", "fvg_batch_upload", "fvg_batch_download",
    "fvg_batch_run", "fvg_batch_set_block", "fvg_batch_get_block", "fvg_batch_tune",
    "fvg_indicator_set", "fvg_indicator_stats", "fvg_indicator_samples", "fvg_trainer_create", "fvg_trainer_destroy",
    "fvg_trainer_last_error", "fvg_trainer_set_data", "fvg_trainer_grad", "fvg_trainer_fit", "fvg_mlp_logits",
    "fvg_checkpoint_save", "fvg_checkpoint_load", "fvg_checkpoint_write", "fvg_checkpoint_read", "fvg_monitor",
    "fvg_conserved_totals_noalloc", "fvg_batch_create_2d", "fvg_batch_set_teams", "fvg_batch_get_plan",
]

MLP_NPARAM = 1601
_F = C.POINTER(C.c_float)


class FvgError(RuntimeError):
    code = -1


class LogicError(FvgError):
    code = ELOGIC


class Diverged(FvgError):
    code = EDIVERGED


class HipError(FvgError):
    code = EHIP


class OutOfMemory(FvgError):
    code = ENOMEM


class Unsupported(FvgError):
    code = EUNSUPPORTED


class InvalidArgument(FvgError, ValueError):
    code = EINVAL


class CheckpointIOError(FvgError, OSError):
    code = EIO


_ERR = {EINVAL: InvalidArgument, ELOGIC: LogicError, EDIVERGED: Diverged, EHIP: HipError, ENOMEM: OutOfMemory,
        EUNSUPPORTED: Unsupported, EIO: CheckpointIOError}

_lib = None


def lib():
    """The loaded library (built by fenice_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(Config)
        L.fvg_config_defaults.argtypes = [P]
        L.fvg_config_set.argtypes = [P, C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t]
        L.fvg_config_finalize.argtypes = [P, C.c_char_p, C.c_size_t]
        L.fvg_dt_length.argtypes = [P]
        L.fvg_dt_length.restype = C.c_double
        L.fvg_create.argtypes = [P, C.c_int, C.POINTER(C.c_void_p)]
        L.fvg_destroy.argtypes = [C.c_void_p]
        L.fvg_last_error.argtypes = [C.c_void_p]
        L.fvg_last_error.restype = C.c_char_p
        L.fvg_upload.argtypes = [C.c_void_p, _D, C.c_size_t, C.c_int]
        L.fvg_download.argtypes = [C.c_void_p, _D, C.c_size_t, C.c_int]
        L.fvg_residual.argtypes = [C.c_void_p, C.c_double, _D, C.c_int, _D]
        L.fvg_step.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, _D, _D]
        L.fvg_advance.argtypes = [C.c_void_p, C.c_int, _D, C.c_double, C.c_double, _I, _D]
        L.fvg_get_lagged_speed.argtypes = [C.c_void_p, _D]
        L.fvg_set_lagged_speed.argtypes = [C.c_void_p, C.c_double]
        L.fvg_conserved_totals.argtypes = [C.c_void_p, _D]
        L.fvg_kernel_times.argtypes = [C.c_void_p, C.POINTER(C.c_char_p), _D, C.POINTER(C.c_longlong), C.c_int]
        L.fvg_set_timing.argtypes = [C.c_void_p, C.c_int]
        L.fvg_set_graphs.argtypes = [C.c_void_p, C.c_int]
        L.fvg_set_fused.argtypes = [C.c_void_p, C.c_int]
        L.fvg_riemann_batch.argtypes = [P, C.c_int, _D, _D, _I, C.c_size_t, _D, _D]
        L.fvg_batch_create.argtypes = [P, C.c_int, _I, C.c_int, C.POINTER(C.c_void_p)]
        L.fvg_batch_create_2d.argtypes = [P, C.c_int, _I, _I, C.c_int, C.POINTER(C.c_void_p)]
        L.fvg_batch_destroy.argtypes = [C.c_void_p]
        L.fvg_batch_last_error.argtypes = [C.c_void_p]
        L.fvg_batch_last_error.restype = C.c_char_p
        L.fvg_batch_upload.argtypes = [C.c_void_p, _D, C.c_longlong, C.c_int]
        L.fvg_batch_download.argtypes = [C.c_void_p, _D, C.c_longlong, C.c_int]
        L.fvg_batch_run.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _D, _I, _I, _D, _D]
        L.fvg_batch_set_block.argtypes = [C.c_void_p, C.c_int]
        L.fvg_batch_get_block.argtypes = [C.c_void_p, _I]
        L.fvg_batch_tune.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_char_p, _I, _I]
        L.fvg_batch_set_teams.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.fvg_batch_get_plan.argtypes = [C.c_void_p, _I, _I, _I]
        F = C.POINTER(C.c_float)
        L.fvg_indicator_set.argtypes = [C.c_void_p, F, C.c_int, C.c_float]
        L.fvg_indicator_stats.argtypes = [C.c_void_p, C.POINTER(C.c_ulonglong), _I, _I]
        L.fvg_indicator_samples.argtypes = [C.c_void_p, C.c_int, F, F, C.c_longlong]
        L.fvg_trainer_create.argtypes = [C.c_int, C.c_longlong, C.POINTER(C.c_void_p)]
        L.fvg_trainer_destroy.argtypes = [C.c_void_p]
        L.fvg_trainer_last_error.argtypes = [C.c_void_p]
        L.fvg_trainer_last_error.restype = C.c_char_p
        L.fvg_trainer_set_data.argtypes = [C.c_void_p, F, F, C.c_longlong]
        L.fvg_trainer_grad.argtypes = [C.c_void_p, F, C.c_float, F, _D]
        L.fvg_trainer_fit.argtypes = [C.c_void_p, F, C.c_int, C.c_int, C.c_float, C.c_float, C.c_ulonglong, _D]
        L.fvg_mlp_logits.argtypes = [C.c_int, F, C.c_float, F, C.c_longlong, F]
        L.fvg_checkpoint_save.argtypes = [P, _D, C.c_size_t, C.c_int, C.c_double, C.c_int, C.c_double, C.c_char_p, C.c_char_p,
                                          C.c_size_t]
        L.fvg_checkpoint_load.argtypes = [P, C.c_char_p, _D, C.c_size_t, C.c_int, _D, _I, _D, C.c_char_p, C.c_size_t]
        L.fvg_checkpoint_write.argtypes = [C.c_void_p, C.c_char_p, C.c_double, C.c_int]
        L.fvg_checkpoint_read.argtypes = [C.c_void_p, C.c_char_p, _D, _I]
        L.fvg_monitor.argtypes = [C.c_void_p, _D]
        L.fvg_conserved_totals_noalloc.argtypes = [C.c_int, C.c_void_p, C.c_longlong, C.c_int, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _check(rc: int, ctx=None):
    if rc == OK:
        return
    msg = lib().fvg_last_error(ctx).decode(errors="replace")
    raise _ERR.get(rc, FvgError)(msg or f"status {rc}")


def config_from_keys(keys: dict) -> Config:
    """A Config from the reference's run keys and labels (fvg_config_set), with `auto` resolved."""
    L = lib()
    c = Config()
    L.fvg_config_defaults(C.byref(c))
    err = C.create_string_buffer(512)
    for k, v in keys.items():
        if isinstance(v, bool):
            v = "true" if v else "false"
        rc = L.fvg_config_set(C.byref(c), str(k).encode(), str(v).encode(), err, 512)
        if rc != OK:
            raise _ERR.get(rc, InvalidArgument)(err.value.decode())
    rc = L.fvg_config_finalize(C.byref(c), err, 512)
    if rc != OK:
        raise InvalidArgument(err.value.decode())
    return c


@dataclass
class StepResult:
    """TimeScheme.hpp:17-20"""
    dt: float
    s_max: float


def _dp(a):
    return a.ctypes.data_as(_D)


class Solver:
    """One device context: the resident state of a uniform mesh and the configured residual and time scheme."""

    def __init__(self, cfg: Config, device: int = 0):
        self.cfg = cfg.copy()
        h = C.c_void_p()
        _check(lib().fvg_create(C.byref(self.cfg), device, C.byref(h)))
        self._h = h
        self.ncells = cfg.nx * cfg.ny * cfg.nz

    def close(self):
        if getattr(self, "_h", None):
            lib().fvg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        _check(rc, self._h)

    @property
    def dt_length(self) -> float:
        return lib().fvg_dt_length(C.byref(self.cfg))

    def upload(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.ndim != 2 or U.shape[0] != self.ncells:
            raise InvalidArgument(f"U must be ({self.ncells}, 4|5)")
        self._c(lib().fvg_upload(self._h, _dp(U), U.shape[0], U.shape[1]))

    def download(self, ncomp: int = 5):
        U = np.empty((self.ncells, ncomp))
        self._c(lib().fvg_download(self._h, _dp(U), self.ncells, ncomp))
        return U

    def residual(self, predictor_half_dt: float = 0.0, ncomp: int = 5):
        """(R, S_max) of the resident state (Simulation::residual)."""
        R = np.empty((self.ncells, ncomp))
        S = C.c_double()
        self._c(lib().fvg_residual(self._h, predictor_half_dt, _dp(R), ncomp, C.byref(S)))
        return R, S.value

    def step(self, t: float, t_max: float, dt_max: float = 0.0) -> StepResult:
        dt, S = C.c_double(), C.c_double()
        self._c(lib().fvg_step(self._h, t, t_max, dt_max, C.byref(dt), C.byref(S)))
        return StepResult(dt.value, S.value)

    def advance(self, t: float, t_end: float, max_steps: int = 0, dt_max: float = 0.0):
        """Steps while t < t_end (at most max_steps > 0); returns (t, steps, s_last)."""
        tt = C.c_double(t)
        n = C.c_int()
        s = C.c_double()
        self._c(lib().fvg_advance(self._h, max_steps, C.byref(tt), t_end, dt_max, C.byref(n), C.byref(s)))
        return tt.value, n.value, s.value

    @property
    def lagged_speed(self) -> float:
        s = C.c_double()
        self._c(lib().fvg_get_lagged_speed(self._h, C.byref(s)))
        return s.value

    @lagged_speed.setter
    def lagged_speed(self, s: float):
        self._c(lib().fvg_set_lagged_speed(self._h, s))

    def checkpoint_write(self, path: str, t: float, step: int):
        """Write the state (and the MUSCL-Hancock lagged speed) as the reference's restart file (fvg_checkpoint_write)."""
        self._c(lib().fvg_checkpoint_write(self._h, os.fsencode(path), t, step))

    def checkpoint_read(self, path: str):
        """Load a restart file (the reference's or checkpoint_write's) into the context; returns (t, step)."""
        t, n = C.c_double(), C.c_int()
        self._c(lib().fvg_checkpoint_read(self._h, os.fsencode(path), C.byref(t), C.byref(n)))
        return t.value, n.value

    def monitor(self):
        """(mass, kinetic_energy, total_energy, res_rho, res_mom, res_E) of the state: a row of the reference's monitor
        file (fvg_monitor)."""
        out = np.empty(6)
        self._c(lib().fvg_monitor(self._h, _dp(out)))
        return out

    def conserved_totals(self):
        out = np.empty(3)
        self._c(lib().fvg_conserved_totals(self._h, _dp(out)))
        return out

    def set_timing(self, on: bool):
        self._c(lib().fvg_set_timing(self._h, int(on)))

    def set_indicator(self, params, threshold: float = 0.0):
        """Weights of the learned troubled-face indicator (FVG_MLP_NPARAM floats); resets its audit counters."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._c(lib().fvg_indicator_set(self._h, p.ctypes.data_as(_F), int(p.size), float(threshold)))

    def indicator_stats(self):
        """(last audit TP FP FN TN, all audits TP FP FN TN, off flag, audits)."""
        c = (C.c_ulonglong * 8)()
        off, n = C.c_int(), C.c_int()
        self._c(lib().fvg_indicator_stats(self._h, c, C.byref(off), C.byref(n)))
        v = [int(x) for x in c]
        return v[:4], v[4:], bool(off.value), n.value

    def samples(self, axis: int):
        """Features (nfaces, 16) and label values (nfaces,) of the faces normal to `axis` of the resident state."""
        cfg = self.cfg
        nf = [(cfg.nx + 1) * cfg.ny * cfg.nz, cfg.nx * (cfg.ny + 1) * cfg.nz, cfg.nx * cfg.ny * (cfg.nz + 1)][axis]
        X = np.empty((nf, 16), dtype=np.float32)
        d = np.empty(nf, dtype=np.float32)
        self._c(lib().fvg_indicator_samples(self._h, axis, X.ctypes.data_as(_F), d.ctypes.data_as(_F), nf))
        return X, d

    def set_fused(self, on: bool):
        """Runge-Kutta stages 2-3: the last axis' faces fused with the update (on) or separate passes (off, the
        default). Same bits."""
        self._c(lib().fvg_set_fused(self._h, int(on)))

    def set_graphs(self, on: bool):
        """advance() replays a captured HIP graph of steps (on, the default) or enqueues every kernel (off)."""
        self._c(lib().fvg_set_graphs(self._h, int(on)))

    def kernel_times(self):
        names = (C.c_char_p * 16)()
        ms = (C.c_double * 16)()
        n = (C.c_longlong * 16)()
        k = lib().fvg_kernel_times(self._h, names, ms, n, 16)
        return {names[i].decode(): (ms[i], n[i]) for i in range(k)}


def riemann_batch(cfg: Config, UL, UR, axis, device: int = 0):
    """Device Riemann fluxes of independent faces (the solver of cfg): (F, smax)."""
    UL = np.ascontiguousarray(UL, np.float64)
    UR = np.ascontiguousarray(UR, np.float64)
    ax = np.ascontiguousarray(axis, np.int32)
    n = UL.shape[0]
    F = np.empty((n, 5))
    S = np.empty(n)
    _check(lib().fvg_riemann_batch(C.byref(cfg), device, _dp(UL), _dp(UR), ax.ctypes.data_as(_I), n, _dp(F), _dp(S)))
    return F, S


@dataclass
class EnsembleResult:
    t: np.ndarray          # final time per instance
    steps: np.ndarray      # steps taken per instance
    diverged: np.ndarray   # 1 where check_progress failed on the instance's last step
    s_last: np.ndarray     # last step's wave speed per instance
    ms: float              # kernel time of the launch


class Ensemble:
    """Independent 1-D runs (instance i: the reference's run on an ncells[i] x 1 mesh) advanced together by one
    kernel launch (fvg_batch_*). `cfg` gives the scheme, x boundaries, Lx and Ly shared by every instance; its
    ny must be 1 with periodic y sides."""

    def __init__(self, cfg: Config, ncells, device: int = 0, ny=None):
        """ncells: cells per instance (N x 1 meshes); with ny, ncells are the nx of 2-D instances nx[i] x ny[i]."""
        if ny is None:
            self.n = np.ascontiguousarray(ncells, dtype=np.int32)
        else:
            self.nx = np.ascontiguousarray(ncells, dtype=np.int32)
            self.ny = np.ascontiguousarray(ny, dtype=np.int32)
            self.n = (self.nx * self.ny).astype(np.int32)
        self.total = int(self.n.sum())
        h = C.c_void_p()
        if ny is None:
            rc = lib().fvg_batch_create(C.byref(cfg), int(self.n.size), self.n.ctypes.data_as(_I), device, C.byref(h))
        else:
            rc = lib().fvg_batch_create_2d(C.byref(cfg), int(self.n.size), self.nx.ctypes.data_as(_I),
                                           self.ny.ctypes.data_as(_I), device, C.byref(h))
        if rc != OK:
            raise _ERR.get(rc, FvgError)(lib().fvg_batch_last_error(None).decode(errors="replace"))
        self._h = h
        self.offsets = np.concatenate([[0], np.cumsum(self.n)]).astype(np.int64)

    def close(self):
        if getattr(self, "_h", None):
            lib().fvg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        if rc != OK:
            raise _ERR.get(rc, FvgError)(lib().fvg_batch_last_error(self._h).decode(errors="replace"))

    def upload(self, U):
        """U: (sum ncells, 3 | 4 | 5) conservative states, instances concatenated in order."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        self._c(lib().fvg_batch_upload(self._h, _dp(U), U.shape[0], U.shape[1]))

    def download(self, ncomp: int = 4):
        U = np.empty((self.total, ncomp))
        self._c(lib().fvg_batch_download(self._h, _dp(U), self.total, ncomp))
        return U

    def instance(self, U, i):
        return U[self.offsets[i]:self.offsets[i + 1]]

    def run(self, t_end: float, max_steps: int, dt_max: float = 0.0) -> EnsembleResult:
        k = self.n.size
        t = np.empty(k)
        steps = np.empty(k, dtype=np.int32)
        div = np.empty(k, dtype=np.int32)
        s = np.empty(k)
        ms = C.c_double()
        self._c(lib().fvg_batch_run(self._h, t_end, dt_max, max_steps, _dp(t), steps.ctypes.data_as(_I),
                                    div.ctypes.data_as(_I), _dp(s), C.byref(ms)))
        return EnsembleResult(t, steps, div, s, ms.value)

    @property
    def block(self) -> int:
        b = C.c_int()
        self._c(lib().fvg_batch_get_block(self._h, C.byref(b)))
        return b.value

    @block.setter
    def block(self, b: int):
        self._c(lib().fvg_batch_set_block(self._h, int(b)))

    def set_teams(self, max_team: int, min_chunk: int = 512, force: bool = False):
        """Teams of workgroups for the largest 1-D instances (max_team = 1: one workgroup per instance)."""
        self._c(lib().fvg_batch_set_teams(self._h, int(max_team), int(min_chunk), 1 if force else 0))

    @property
    def plan(self):
        """(teams, workgroups in teams, workgroups launched) of the last run."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._c(lib().fvg_batch_get_plan(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def tune(self, t_probe: float, max_steps: int, cache_path: str | None = None):
        """Autotune the workgroup size on a probe run (the state is restored); returns (block, from_cache)."""
        b, fc = C.c_int(), C.c_int()
        self._c(lib().fvg_batch_tune(self._h, t_probe, max_steps, cache_path.encode() if cache_path else None,
                                     C.byref(b), C.byref(fc)))
        return b.value, bool(fc.value)


class Trainer:
    """Training of the indicator network on the device (fvg_trainer_*)."""

    def __init__(self, capacity: int, device: int = 0):
        h = C.c_void_p()
        rc = lib().fvg_trainer_create(device, int(capacity), C.byref(h))
        if rc != OK:
            raise _ERR.get(rc, FvgError)(lib().fvg_trainer_last_error(None).decode(errors="replace"))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().fvg_trainer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        if rc != OK:
            raise _ERR.get(rc, FvgError)(lib().fvg_trainer_last_error(self._h).decode(errors="replace"))

    def set_data(self, X, y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        if X.shape[1] == 15:
            X = np.ascontiguousarray(np.concatenate([X, np.ones((X.shape[0], 1), np.float32)], axis=1))
        y = np.ascontiguousarray(y, dtype=np.float32)
        self._keep = (X, y)
        self._c(lib().fvg_trainer_set_data(self._h, X.ctypes.data_as(_F), y.ctypes.data_as(_F), X.shape[0]))

    def grad(self, params, pos_weight: float = 1.0):
        p = np.ascontiguousarray(params, dtype=np.float32)
        g = np.empty(MLP_NPARAM, dtype=np.float32)
        loss = C.c_double()
        self._c(lib().fvg_trainer_grad(self._h, p.ctypes.data_as(_F), pos_weight, g.ctypes.data_as(_F), C.byref(loss)))
        return loss.value, g

    def fit(self, params, epochs: int, batch: int, lr: float, pos_weight: float = 1.0, seed: int = 0):
        p = np.array(params, dtype=np.float32)
        loss = C.c_double()
        self._c(lib().fvg_trainer_fit(self._h, p.ctypes.data_as(_F), epochs, batch, lr, pos_weight, seed, C.byref(loss)))
        return p, loss.value


def mlp_logits(params, X, threshold: float = 0.0, device: int = 0):
    """Logits of feature rows X (n, 15 or 16) through the face kernels' inference routine."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    if X.shape[1] == 15:
        X = np.ascontiguousarray(np.concatenate([X, np.ones((X.shape[0], 1), np.float32)], axis=1))
    p = np.ascontiguousarray(params, dtype=np.float32)
    z = np.empty(X.shape[0], dtype=np.float32)
    rc = lib().fvg_mlp_logits(device, p.ctypes.data_as(_F), threshold, X.ctypes.data_as(_F), X.shape[0], z.ctypes.data_as(_F))
    if rc != OK:
        raise _ERR.get(rc, FvgError)(lib().fvg_trainer_last_error(None).decode(errors="replace"))
    return z


def _host_err(rc: int, err) -> None:
    if rc != OK:
        raise _ERR.get(rc, FvgError)(err.value.decode(errors="replace") or f"status {rc}")


def checkpoint_save(cfg: Config, U, t: float, step: int, path: str, lagged_speed: float = 0.0) -> None:
    """Write U (ncells x 4 in 2-D, x 5 in 3-D) with t, step and lagged_speed as a restart file (host only,
    fvg_checkpoint_save)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    err = C.create_string_buffer(512)
    _host_err(lib().fvg_checkpoint_save(C.byref(cfg), _dp(U), U.shape[0], U.shape[1], t, step, lagged_speed, os.fsencode(path),
                                        err, 512), err)


def checkpoint_load(cfg: Config, path: str, ncomp: int | None = None):
    """Read a restart file of cfg's mesh (host only, fvg_checkpoint_load): (U, t, step, lagged_speed)."""
    nc = cfg.nx * cfg.ny * cfg.nz
    nq = ncomp or (5 if cfg.nz > 1 else 4)
    U = np.empty((nc, nq))
    t, n, s = C.c_double(), C.c_int(), C.c_double()
    err = C.create_string_buffer(512)
    _host_err(lib().fvg_checkpoint_load(C.byref(cfg), os.fsencode(path), _dp(U), nc, nq, C.byref(t), C.byref(n), C.byref(s), err, 512),
              err)
    return U, t.value, n.value, s.value
