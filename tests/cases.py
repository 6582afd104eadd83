"""Configurations shared by the parity tests: reference run keys (labels as in the case files) and the
seeded states they are checked on."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
TWO_PI = "6.283185307179586"

PERIODIC2 = dict(bc_xmin="periodic", bc_xmax="periodic", bc_ymin="periodic", bc_ymax="periodic")
PERIODIC3 = dict(PERIODIC2, bc_zmin="periodic", bc_zmax="periodic")


def golden_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load_golden(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    return dict(U0=d["U0"], U_ref=d["U_ref"], step=int(d["step"]), keys=json.loads(str(d["keys"])), case=str(d["case"]))


def smooth_state(nx, ny, nz, seed, gamma=1.4, noise=0.02, mach=0.3, p_inf=0.0):
    """A smooth, non-trivial conservative state (n, 5) with seeded noise; rho*w = +0 in 2-D. p_inf > 0: the energy
    of the stiffened gas, (p + gamma p_inf) / (gamma - 1) + kinetic (Core.hpp:209-215)."""
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x = (i.ravel() + 0.5) / nx * 2 * np.pi
    y = (j.ravel() + 0.5) / ny * 2 * np.pi
    z = (k.ravel() + 0.5) / max(nz, 1) * 2 * np.pi
    n = nx * ny * nz
    rho = 1.0 + 0.2 * np.sin(x) * np.cos(y) + noise * rng.standard_normal(n)
    u = mach * (np.sin(y + 0.3) + noise * rng.standard_normal(n))
    v = mach * (np.cos(x) * np.sin(z) + noise * rng.standard_normal(n))
    w = mach * (np.sin(x + y) + noise * rng.standard_normal(n)) if nz > 1 else np.zeros(n)
    p = 1.0 + 0.1 * np.cos(x + 2 * y) + noise * np.abs(rng.standard_normal(n))
    U = np.empty((n, 5))
    U[:, 0] = rho
    U[:, 1] = rho * u
    U[:, 2] = rho * v
    U[:, 3] = (p + gamma * p_inf) / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    U[:, 4] = rho * w if nz > 1 else 0.0
    return U


def shock_state(nx, ny, nz, seed, gamma=1.4, p_inf=0.0):
    """A state with strong jumps (pressure ratio 10, density ratio 8) plus seeded noise: exercises the shock blend,
    the limiters' clipping and the Hancock admissibility reverts."""
    U = smooth_state(nx, ny, nz, seed, gamma, noise=0.01, mach=0.5, p_inf=p_inf)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    left = (i.ravel() + j.ravel()) < (nx + ny) // 2
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    U[:, 0] *= rho
    U[:, 1] *= rho
    U[:, 2] *= rho
    U[:, 4] *= rho
    kin = 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2 + U[:, 4] ** 2) / U[:, 0]
    U[:, 3] = (p + gamma * p_inf) / (gamma - 1.0) + kin
    return U


def base_keys(nx, ny, nz=1, periodic=True, **kw):
    k = dict(nx=nx, ny=ny, nz=nz, Lx=TWO_PI, Ly=TWO_PI, Lz=TWO_PI, gamma=1.4, use_viscous="false", riemann_solver="HLLC",
             reconstruction="muscl", muscl_limiter="Van Leer", time_scheme="SSP Runge-Kutta 3", cfl=0.5, boundary_order=1)
    if periodic:
        k.update(PERIODIC3 if nz > 1 else PERIODIC2)
    k.update(kw)
    return {a: str(b) for a, b in k.items()}


# (id, keys, state kind): the configuration matrix of the parity tests
def matrix():
    out = []
    for lim in ("Minmod", "Van Leer", "MC", "Koren"):
        out.append((f"2d-muscl-{lim}", base_keys(12, 10, muscl_limiter=lim), "shock"))
    for rc in ("weno3", "weno5", "none"):
        out.append((f"2d-{rc}", base_keys(12, 10, reconstruction=rc, time_scheme="SSP Runge-Kutta 3" if rc != "none" else "Explicit Euler"), "shock"))
    out.append(("2d-hancock", base_keys(14, 12, time_scheme="MUSCL-Hancock"), "shock"))
    out.append(("2d-lowmach-rk2", base_keys(12, 10, riemann_solver="HLLC (Low-Mach Thornber)", time_scheme="SSP Runge-Kutta 2"), "smooth"))
    out.append(("2d-visc-periodic", base_keys(12, 10, use_viscous="true", mu="0.02", time_scheme="SSP Runge-Kutta 2"), "smooth"))
    walls = dict(bc_xmin="transmissive", bc_xmax="wall", bc_ymin="no_slip", bc_ymax="wall")
    for bo in (1, 2):
        out.append((f"2d-walls-muscl-bo{bo}", base_keys(12, 10, periodic=False, boundary_order=bo, **walls), "shock"))
        out.append((f"2d-walls-weno5-bo{bo}", base_keys(12, 10, periodic=False, boundary_order=bo, reconstruction="weno5", **walls), "shock"))
        out.append((f"2d-walls-hancock-bo{bo}", base_keys(12, 10, periodic=False, boundary_order=bo, time_scheme="MUSCL-Hancock", **walls), "shock"))
        out.append((f"2d-walls-visc-bo{bo}", base_keys(12, 10, periodic=False, boundary_order=bo, use_viscous="true", mu="0.01",
                                                       **dict(walls, bc_ymax="moving_wall", u_in="1.0")), "smooth"))
    out.append(("2d-first-order-walls", base_keys(12, 10, periodic=False, reconstruction="none", time_scheme="Explicit Euler", **walls), "shock"))
    for rc in ("weno5", "weno3", "muscl"):
        out.append((f"3d-{rc}-visc-lm", base_keys(8, 9, 7, reconstruction=rc, use_viscous="true", mu="0.01", pr="0.71",
                                                    riemann_solver="HLLC (Low-Mach Thornber)"), "smooth"))
    out.append(("3d-weno5-hllc-shock", base_keys(8, 9, 7, reconstruction="weno5"), "shock"))
    out.append(("3d-hancock", base_keys(8, 9, 7, time_scheme="MUSCL-Hancock"), "shock"))
    out.append(("3d-hancock-visc", base_keys(8, 9, 7, time_scheme="MUSCL-Hancock", use_viscous="true", mu="0.01"), "smooth"))
    walls3 = dict(bc_xmin="transmissive", bc_xmax="wall", bc_ymin="no_slip", bc_ymax="wall", bc_zmin="wall", bc_zmax="no_slip")
    for bo in (1, 2):
        out.append((f"3d-walls-weno5-visc-bo{bo}", base_keys(8, 9, 7, periodic=False, boundary_order=bo, reconstruction="weno5",
                                                             use_viscous="true", mu="0.01", **walls3), "smooth"))
        out.append((f"3d-walls-hancock-bo{bo}", base_keys(8, 9, 7, periodic=False, boundary_order=bo, time_scheme="MUSCL-Hancock",
                                                          **walls3), "shock"))
    # thin meshes: axes as deep as the stencil's halo (periodic wrap or ghost slab of exactly Dh cells) or shallower
    # (no halo: edge copies), where every halo layer comes from a cell next to the other side or from the edge cell
    out.append(("3d-thin-weno5-periodic-n3", base_keys(5, 3, 3, reconstruction="weno5"), "shock"))
    out.append(("3d-thin-weno5-periodic-n2", base_keys(6, 4, 2, reconstruction="weno5"), "shock"))
    out.append(("3d-thin-visc-weno5-n3", base_keys(4, 3, 5, reconstruction="weno5", use_viscous="true", mu="0.01"), "smooth"))
    out.append(("3d-thin-muscl-walls-bo2-n2", base_keys(5, 2, 3, periodic=False, boundary_order=2, **walls3), "shock"))
    out.append(("2d-thin-weno5-walls-bo2-n3", base_keys(7, 3, periodic=False, boundary_order=2, reconstruction="weno5", **walls), "shock"))
    out.append(("2d-thin-weno3-periodic-n2", base_keys(9, 2, reconstruction="weno3"), "shock"))
    out.append(("2d-thin-hancock-walls-bo2-n2", base_keys(6, 2, periodic=False, boundary_order=2, time_scheme="MUSCL-Hancock", **walls),
                "shock"))
    # a mass-flow inlet (the side's mean density sets the inflow velocity) and a non-reflecting outlet (characteristic
    # ghost relaxed towards p_out, transverse terms), in 2-D and 3-D, every scheme family
    io = dict(bc_xmin="massflow_inlet", bc_xmax="nonreflecting", bc_ymin="wall", bc_ymax="no_slip", u_in="0.5", v_in="0.1",
              rho_in="1.2", p_out="0.9")
    io3 = dict(io, bc_zmin="periodic", bc_zmax="periodic")
    for bo in (1, 2):
        out.append((f"2d-inflow-nrbc-muscl-bo{bo}", base_keys(12, 10, periodic=False, boundary_order=bo, **io), "shock"))
    out.append(("2d-inflow-nrbc-weno5", base_keys(12, 10, periodic=False, boundary_order=2, reconstruction="weno5", **io), "smooth"))
    out.append(("2d-inflow-nrbc-hancock", base_keys(12, 10, periodic=False, boundary_order=2, time_scheme="MUSCL-Hancock", **io), "shock"))
    out.append(("2d-inflow-nrbc-visc", base_keys(12, 10, periodic=False, boundary_order=2, use_viscous="true", mu="0.01", **io), "smooth"))
    out.append(("2d-nrbc-beta-sigma", base_keys(12, 10, periodic=False, boundary_order=2, nrbc_beta="0.5", nrbc_sigma="0.6", **io), "shock"))
    out.append(("2d-inflow-nrbc-y-periodic-x", base_keys(12, 10, periodic=False, bc_xmin="periodic", bc_xmax="periodic",
                                                         bc_ymin="massflow_inlet", bc_ymax="nonreflecting", u_in="0.3", rho_in="0.9",
                                                         p_out="1.1"), "smooth"))
    out.append(("3d-inflow-nrbc-weno5-visc", base_keys(8, 9, 7, periodic=False, boundary_order=2, reconstruction="weno5", use_viscous="true",
                                                       mu="0.01", **io3), "smooth"))
    out.append(("3d-inflow-nrbc-z-walls", base_keys(8, 9, 7, periodic=False, boundary_order=1, **dict(io, bc_zmin="wall",
                                                                                                       bc_zmax="nonreflecting")), "smooth"))
    out.append(("3d-inflow-nrbc-hancock", base_keys(8, 9, 7, periodic=False, boundary_order=2, time_scheme="MUSCL-Hancock", **io3), "shock"))
    # a velocity inlet and a pressure outlet (velocity kept where the flow leaves, zeroed where it enters; pressure
    # p_out), with a transmissive side and a moving wall, in 2-D and 3-D
    vo = dict(bc_xmin="vel_inlet", bc_xmax="outlet", bc_ymin="transmissive", bc_ymax="moving_wall", u_in="0.6", v_in="0.1",
              p_out="0.8")
    vo3 = dict(vo, bc_zmin="outlet", bc_zmax="vel_inlet")
    for bo in (1, 2):
        out.append((f"2d-velin-outlet-muscl-bo{bo}", base_keys(12, 10, periodic=False, boundary_order=bo, **vo), "shock"))
    out.append(("2d-velin-outlet-weno5", base_keys(12, 10, periodic=False, boundary_order=2, reconstruction="weno5", **vo), "shock"))
    out.append(("2d-velin-outlet-hancock", base_keys(12, 10, periodic=False, boundary_order=2, time_scheme="MUSCL-Hancock", **vo), "shock"))
    out.append(("2d-velin-outlet-visc", base_keys(12, 10, periodic=False, boundary_order=2, use_viscous="true", mu="0.01", **vo), "smooth"))
    out.append(("3d-velin-outlet-weno5-visc", base_keys(8, 9, 7, periodic=False, boundary_order=2, reconstruction="weno5", use_viscous="true",
                                                        mu="0.01", **vo3), "smooth"))
    out.append(("3d-velin-outlet-hancock-bo1", base_keys(8, 9, 7, periodic=False, boundary_order=1, time_scheme="MUSCL-Hancock", **vo3), "shock"))
    # the stiffened gas (the reference's sod_stiffened_grav values, gamma 4.4 and p_inf 0.5): every scheme family,
    # walls with ghost slabs, a non-reflecting outlet and the viscous term
    st = dict(eos_type="Stiffened Gas", gamma="4.4", p_inf="0.5")
    out.append(("2d-stiffened-muscl-walls-bo2", base_keys(12, 10, periodic=False, boundary_order=2, muscl_limiter="MC", **walls, **st),
                "shock"))
    out.append(("2d-stiffened-weno5-lowmach", base_keys(12, 10, reconstruction="weno5", riemann_solver="HLLC (Low-Mach Thornber)", **st),
                "shock"))
    out.append(("2d-stiffened-hancock", base_keys(14, 12, time_scheme="MUSCL-Hancock", **st), "shock"))
    out.append(("2d-stiffened-inflow-nrbc", base_keys(12, 10, periodic=False, boundary_order=2, **io, **st), "smooth"))
    out.append(("3d-stiffened-weno5-visc", base_keys(8, 9, 7, reconstruction="weno5", use_viscous="true", mu="0.01", pr="0.71", **st),
                "smooth"))
    out.append(("3d-stiffened-walls-weno3-bo2", base_keys(8, 9, 7, periodic=False, boundary_order=2, reconstruction="weno3", **walls3,
                                                          **st), "shock"))
    return out


# flux_quadrature = fourth: WENO3-Z / WENO5-Z of the conserved averages with the transverse correction, over periodic
# and wall boundaries (ghost slabs with boundary_order 2, first-order side faces with 1), viscous and low-Mach, the
# shock sensor on and off, and a periodic axis shallower than the halo (no halo, yet the corrections wrap)
def fourth_matrix():
    F = dict(flux_quadrature="fourth")
    walls = dict(bc_xmin="transmissive", bc_xmax="wall", bc_ymin="no_slip", bc_ymax="wall")
    walls3 = dict(walls, bc_zmin="wall", bc_zmax="no_slip")
    return [
        ("2d-weno5-fourth", base_keys(12, 10, reconstruction="weno5", **F), "shock"),
        ("2d-weno5-fourth-noblend", base_keys(12, 10, reconstruction="weno5", shock_blend="0", **F), "shock"),
        ("2d-weno3-fourth-smooth", base_keys(12, 10, reconstruction="weno3", **F), "smooth"),
        ("2d-walls-weno5-fourth-bo2", base_keys(12, 10, periodic=False, boundary_order=2, reconstruction="weno5", **walls, **F), "shock"),
        ("2d-walls-weno3-fourth-bo1", base_keys(12, 10, periodic=False, boundary_order=1, reconstruction="weno3", **walls, **F), "shock"),
        ("2d-walls-weno5-fourth-visc", base_keys(12, 10, periodic=False, boundary_order=2, reconstruction="weno5", use_viscous="true",
                                                 mu="0.01", **dict(walls, bc_ymax="moving_wall", u_in="1.0"), **F), "smooth"),
        ("3d-weno5-fourth-visc-lm", base_keys(8, 9, 7, reconstruction="weno5", use_viscous="true", mu="0.01",
                                              riemann_solver="HLLC (Low-Mach Thornber)", **F), "smooth"),
        ("3d-weno5-fourth-shock", base_keys(8, 9, 7, reconstruction="weno5", **F), "shock"),
        ("3d-walls-weno3-fourth-bo2", base_keys(8, 9, 7, periodic=False, boundary_order=2, reconstruction="weno3", **walls3, **F), "shock"),
        ("3d-walls-weno5-fourth-bo1", base_keys(8, 9, 7, periodic=False, boundary_order=1, reconstruction="weno5", **walls3, **F), "shock"),
        ("3d-thin-weno5-fourth-n2", base_keys(6, 4, 2, reconstruction="weno5", **F), "shock"),
        ("2d-euler-weno5-fourth", base_keys(12, 10, reconstruction="weno5", time_scheme="Explicit Euler", **F), "smooth"),
        ("3d-inflow-nrbc-weno5-fourth", base_keys(8, 9, 7, periodic=False, boundary_order=2, reconstruction="weno5",
                                                  **dict(walls, bc_xmin="massflow_inlet", bc_xmax="nonreflecting", bc_zmin="periodic",
                                                         bc_zmax="periodic", u_in="0.5", v_in="0.1", rho_in="1.2", p_out="0.9"), **F),
         "shock"),
        ("2d-velin-outlet-weno5-fourth-bo1", base_keys(12, 10, periodic=False, boundary_order=1, reconstruction="weno5",
                                                       **dict(walls, bc_xmin="vel_inlet", bc_xmax="outlet", u_in="0.6", v_in="0.1",
                                                              p_out="0.8"), **F), "shock"),
        ("3d-velin-outlet-weno3-fourth-visc", base_keys(8, 9, 7, periodic=False, boundary_order=2, reconstruction="weno3", use_viscous="true",
                                                        mu="0.01", **dict(walls3, bc_xmin="vel_inlet", bc_xmax="outlet", bc_zmin="outlet",
                                                                          u_in="0.6", v_in="0.1", p_out="0.8"), **F), "smooth"),
    ]


def state_for(kind, keys, seed=7):
    nx, ny, nz = int(keys["nx"]), int(keys["ny"]), int(keys.get("nz", 1))
    if keys.get("eos_type", "Ideal Gas").lower() == "stiffened gas":
        gamma, p_inf = float(keys["gamma"]), float(keys["p_inf"])
        return (shock_state if kind == "shock" else smooth_state)(nx, ny, nz, seed, gamma=gamma, p_inf=p_inf)
    return (shock_state if kind == "shock" else smooth_state)(nx, ny, nz, seed)


# The isentropic vortex on 256^2 under MUSCL-Hancock (Van Leer, HLLC) at CFL 0.8: the reference's run diverges at
# step 65 (dt falls below the resolution of t), so parity covers steps 1-64 and the divergence itself
VORTEX_HANCOCK_256 = dict(nx="256", ny="256", nz="1", Lx="10.0", Ly="10.0", gamma="1.4", use_viscous="false",
                          bc_xmin="periodic", bc_xmax="periodic", bc_ymin="periodic", bc_ymax="periodic",
                          riemann_solver="HLLC", reconstruction="muscl", muscl_limiter="Van Leer",
                          time_scheme="MUSCL-Hancock", cfl="0.8", boundary_order="1")   # = workloads.vortex_keys(256)
VORTEX_DIVERGES_AT = 65


def vortex_state(n, L=10.0, gamma=1.4):
    """The isentropic vortex (fenice_amd.workloads.vortex_state) with the reference's 5-slot State layout."""
    from fenice_amd.workloads import vortex_state as v4
    U = np.zeros((n * n, 5))
    U[:, :4] = v4(n, L, gamma)
    return U


def oracle_instance(oracle, keys, n, U4, t_end, max_steps, dt_max=0.0):
    """One instance of an ensemble run through the oracle as the reference's N x 1 run: steps while t < t_end and
    fewer than max_steps; a failed check_progress ends the run with that step counted. Returns
    (U (n, 4), t, steps, diverged)."""
    from fenice_amd.fvg import config_from_keys
    cfg = config_from_keys(dict(keys, nx=str(n)))
    U = np.zeros((n, 5))
    U[:, :4] = U4
    t, steps, lag, diverged = 0.0, 0, 0.0, 0
    while t < t_end and steps < max_steps:
        dt, _, lag, ok = oracle.step(cfg, U, t, t_end, dt_max, lag)
        steps += 1
        if not ok:
            diverged = 1
            break
        t += dt
    return U[:, :4].copy(), t, steps, diverged


def own_case_file(directory) -> str:
    """A run description of our own for the reference's case loader: a uniform state, every key of the run given as an
    override and the state set explicitly, so a reference run needs none of the reference's case folders."""
    import os
    path = os.path.join(str(directory), "case.toml")
    with open(path, "w") as f:
        f.write('[case]\nname = "Uniform run"\n\n[initial]\nstate = { rho = 1.0, u = 0.0, v = 0.0, p = 1.0 }\n')
    return path
