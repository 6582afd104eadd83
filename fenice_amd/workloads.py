"""Synthetic workloads of the uniform path, generated on the host with documented counter-based randomness so that
every run (device, oracle, reference) starts from the same bits.

- `riemann_ensemble`: independent 1-D Riemann problems on [0, 1] (transmissive ends), instance i with N_i cells,
  N_i ~ U{nmin..nmax}, rho_L, rho_R ~ logU[0.1, 10], p_L, p_R ~ logU[0.01, 1000], u_L, u_R ~ U[-2, 2],
  x0 ~ U[0.3, 0.7], drawn in that order per instance from one SplitMix64 stream (seed 1024 for the 1024-instance
  ensemble).
- `taylor_green`: the 3-D Taylor-Green vortex at Mach 0.1 (the benchmark's workload).
- `kelvin_helmholtz`: the periodic shear layer on [0, 1]^2 (rho 2, u -0.5 for |y - 0.5| < 0.25, else rho 1,
  u 0.5; p 2.5; v = 0.01 sin(4 pi x) + U(-1e-3, 1e-3)), the noise from SplitMix64(seed) indexed by cell.
"""
from __future__ import annotations

import numpy as np

M64 = (1 << 64) - 1
GOLDEN64 = 0x9E3779B97F4A7C15


def splitmix64(seed: int, count: int) -> np.ndarray:
    """`count` outputs of SplitMix64 (Steele, Lea and Flood 2014) started at `seed`, as uint64."""
    s = (np.uint64(seed) + np.arange(1, count + 1, dtype=np.uint64) * np.uint64(GOLDEN64))
    with np.errstate(over="ignore"):
        z = s
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def unit_doubles(seed: int, count: int) -> np.ndarray:
    """Uniform doubles in [0, 1): the top 53 bits of each SplitMix64 output times 2^-53."""
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def riemann_ensemble(ninst: int = 1024, seed: int = 1024, nmin: int = 128, nmax: int = 8192, gamma: float = 1.4):
    """(ncells int32[ninst], U float64[sum ncells, 4] in the reference's 2-D State layout, params float64[ninst, 7]
    = rho_L, rho_R, p_L, p_R, u_L, u_R, x0)."""
    r = unit_doubles(seed, 8 * ninst).reshape(ninst, 8)
    n = (nmin + np.floor(r[:, 0] * (nmax - nmin + 1))).astype(np.int32)
    rho_l = 10.0 ** (-1.0 + 2.0 * r[:, 1])
    rho_r = 10.0 ** (-1.0 + 2.0 * r[:, 2])
    p_l = 10.0 ** (-2.0 + 5.0 * r[:, 3])
    p_r = 10.0 ** (-2.0 + 5.0 * r[:, 4])
    u_l = -2.0 + 4.0 * r[:, 5]
    u_r = -2.0 + 4.0 * r[:, 6]
    x0 = 0.3 + 0.4 * r[:, 7]
    params = np.stack([rho_l, rho_r, p_l, p_r, u_l, u_r, x0], axis=1)
    return n, riemann_states(n, params, gamma), params


def riemann_states(n, params, gamma: float = 1.4) -> np.ndarray:
    """Concatenated (sum n, 4) conservative states of Riemann problems on [0, 1]: cell centre x < x0 takes the left
    state. params[i] = rho_L, rho_R, p_L, p_R, u_L, u_R, x0."""
    U = np.zeros((int(np.sum(n)), 4))
    off = 0
    for i in range(len(n)):
        rho_l, rho_r, p_l, p_r, u_l, u_r, x0 = params[i]
        x = (np.arange(n[i]) + 0.5) / n[i]
        left = x < x0
        rho = np.where(left, rho_l, rho_r)
        u = np.where(left, u_l, u_r)
        p = np.where(left, p_l, p_r)
        s = U[off:off + n[i]]
        s[:, 0] = rho
        s[:, 1] = rho * u
        s[:, 3] = p / (gamma - 1.0) + 0.5 * rho * (u * u)
        off += n[i]
    return U


def ensemble_keys(t_end: float = 0.15) -> dict:
    """Run keys of every instance (nx is set per instance): MUSCL MC + HLLC + SSP-RK2 at CFL 0.5 on [0, 1] with
    transmissive x ends; one periodic row in y."""
    return dict(nx="128", ny="1", nz="1", Lx="1.0", Ly="1.0", gamma="1.4", use_viscous="false", bc_xmin="transmissive",
                bc_xmax="transmissive", bc_ymin="periodic", bc_ymax="periodic", riemann_solver="HLLC",
                reconstruction="muscl", muscl_limiter="MC", time_scheme="SSP Runge-Kutta 2", cfl="0.5",
                t_end=str(t_end))


def kelvin_helmholtz(n: int, seed: int = 20261202, gamma: float = 1.4) -> np.ndarray:
    """(n*n, 4) conservative state of the shear layer on [0, 1]^2, cell (i, j) at row-major index j*n + i."""
    h = 1.0 / n
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = (i.ravel() + 0.5) * h
    y = (j.ravel() + 0.5) * h
    inner = np.abs(y - 0.5) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    u = np.where(inner, -0.5, 0.5)
    noise = -1e-3 + 2e-3 * unit_doubles(seed, n * n)
    v = 0.01 * np.sin(4.0 * np.pi * x) + noise
    p = np.full(n * n, 2.5)
    U = np.empty((n * n, 4))
    U[:, 0] = rho
    U[:, 1] = rho * u
    U[:, 2] = rho * v
    U[:, 3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return U


def kh_keys(n: int) -> dict:
    return dict(nx=str(n), ny=str(n), nz="1", Lx="1.0", Ly="1.0", gamma="1.4", use_viscous="false", bc_xmin="periodic",
                bc_xmax="periodic", bc_ymin="periodic", bc_ymax="periodic", riemann_solver="HLLC", reconstruction="weno5",
                time_scheme="SSP Runge-Kutta 3", cfl="0.5")


def taylor_green(n: int, gamma: float = 1.4, mach: float = 0.1) -> np.ndarray:
    """(n^3, 5) conservative state of the 3-D Taylor-Green vortex on [0, 2 pi]^3 (cpp/src/TestCases.cpp:262-271):
    rho 1, u = sin x cos y cos z, v = -cos x sin y cos z, w = 0,
    p = 1 / (gamma Ma^2) + (cos 2x + cos 2y)(cos 2z + 2) / 16."""
    h = 2.0 * np.pi / n
    x = (np.arange(n) + 0.5) * h
    z3, y3, x3 = np.meshgrid(x, x, x, indexing="ij")
    u = np.sin(x3) * np.cos(y3) * np.cos(z3)
    v = -np.cos(x3) * np.sin(y3) * np.cos(z3)
    p = 1.0 / (gamma * mach * mach) + (np.cos(2 * x3) + np.cos(2 * y3)) * (np.cos(2 * z3) + 2.0) / 16.0
    U = np.empty((n ** 3, 5))
    U[:, 0] = 1.0
    U[:, 1] = u.ravel()
    U[:, 2] = v.ravel()
    U[:, 3] = p.ravel() / (gamma - 1.0) + 0.5 * (u * u + v * v).ravel()
    U[:, 4] = 0.0
    return U


def taylor_green_keys(n: int, flux: str = "fp64", quadrature: str = "midpoint") -> dict:
    """Navier-Stokes at Re 1600 (mu 6.25e-4), Pr 0.71, WENO5-Z + HLLC + SSP-RK3, CFL 0.5, periodic; `quadrature`
    "fourth" takes the fourth-order flux quadrature (FP64)."""
    L = "6.283185307179586"
    k = dict(nx=n, ny=n, nz=n, Lx=L, Ly=L, Lz=L, gamma=1.4, use_viscous="true", mu="6.25e-4", pr="0.71",
             bc_xmin="periodic", bc_xmax="periodic", bc_ymin="periodic", bc_ymax="periodic", bc_zmin="periodic",
             bc_zmax="periodic", riemann_solver="HLLC", reconstruction="weno5", time_scheme="SSP Runge-Kutta 3", cfl="0.5",
             flux_precision=flux)
    if quadrature != "midpoint":
        k["flux_quadrature"] = quadrature
    return k


def vortex_state(n: int, L: float = 10.0, gamma: float = 1.4) -> np.ndarray:
    """(n*n, 4) conservative state of the isentropic vortex (beta 5, free stream (1, 1), centre (L/2, L/2)) on an n x n
    periodic square (cpp/src/TestCases.cpp:284-295)."""
    h = L / n
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    dx = (i.ravel() + 0.5) * h - 0.5 * L
    dy = (j.ravel() + 0.5) * h - 0.5 * L
    dx = dx - L * np.round(dx / L)
    dy = dy - L * np.round(dy / L)
    r2 = dx * dx + dy * dy
    g = 5.0 / (2.0 * np.pi) * np.exp(0.5 * (1.0 - r2))
    T = 1.0 - (gamma - 1.0) * 25.0 / (8.0 * gamma * np.pi * np.pi) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    u, v, p = 1.0 - g * dy, 1.0 + g * dx, rho * T
    U = np.zeros((n * n, 4))
    U[:, 0], U[:, 1], U[:, 2] = rho, rho * u, rho * v
    U[:, 3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return U


def vortex_rect(nx: int, ny: int, L: float = 10.0, gamma: float = 1.4) -> np.ndarray:
    """(nx*ny, 4) conservative state of the isentropic vortex (beta 5, free stream (1, 1), centre (L/2, L/2)) on an
    nx x ny mesh of [0, L]^2 (cpp/src/TestCases.cpp:284-295; vortex_state is the square case)."""
    hx, hy = L / nx, L / ny
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    dx = (i.ravel() + 0.5) * hx - 0.5 * L
    dy = (j.ravel() + 0.5) * hy - 0.5 * L
    r2 = dx * dx + dy * dy
    g = 5.0 / (2.0 * np.pi) * np.exp(0.5 * (1.0 - r2))
    T = 1.0 - (gamma - 1.0) * 25.0 / (8.0 * gamma * np.pi * np.pi) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    u, v, p = 1.0 - g * dy, 1.0 + g * dx, rho * T
    U = np.zeros((nx * ny, 4))
    U[:, 0], U[:, 1], U[:, 2] = rho, rho * u, rho * v
    U[:, 3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return U


def vortex_ensemble(ninst: int = 256, seed: int = 256, nmin: int = 32, nmax: int = 256):
    """Isentropic vortices on square periodic meshes of uneven resolution, n uniform in [nmin, nmax]:
    (n int32[ninst], U float64[sum n^2, 4])."""
    r = unit_doubles(seed, ninst)
    n = (nmin + np.floor(r * (nmax - nmin + 1))).astype(np.int32)
    return n, np.concatenate([vortex_rect(int(k), int(k)) for k in n])


def vortex_ensemble_keys(**over) -> dict:
    """MUSCL (Van Leer) + HLLC + SSP-RK2, CFL 0.5, periodic [0, 10]^2 (the mesh size is each instance's)."""
    k = dict(nx="8", ny="8", nz="1", Lx="10.0", Ly="10.0", gamma="1.4", use_viscous="false", bc_xmin="periodic",
             bc_xmax="periodic", bc_ymin="periodic", bc_ymax="periodic", riemann_solver="HLLC", reconstruction="muscl",
             muscl_limiter="Van Leer", time_scheme="SSP Runge-Kutta 2", cfl="0.5", boundary_order="1")
    k.update(over)
    return k


def vortex_keys(n: int, cfl: str = "0.8", time_scheme: str = "MUSCL-Hancock") -> dict:
    """MUSCL (Van Leer) + HLLC, periodic [0, 10]^2."""
    return dict(nx=str(n), ny=str(n), nz="1", Lx="10.0", Ly="10.0", gamma="1.4", use_viscous="false", bc_xmin="periodic",
                bc_xmax="periodic", bc_ymin="periodic", bc_ymax="periodic", riemann_solver="HLLC", reconstruction="muscl",
                muscl_limiter="Van Leer", time_scheme=time_scheme, cfl=cfl, boundary_order="1")
