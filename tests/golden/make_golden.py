"""Writes the golden fixtures of the uniform-grid path (run here, where the reference sources are; the
fixtures travel with the repository).

For each regression case of the reference's cpp/tests/regression.py:43-237 that lies on the path, the fixture
holds
  U0     the case's initial field, made by the reference itself (Simulation::initialize_field of the
         reference built from its sources by oracle/Makefile), so that parity runs start from the same bits;
  U_ref  the reference's own stored final state, cpp/tests/reference/<name>.npz (data, copied as is);
  step   the stored step count;
  keys   the run keys: the case file's values (cases/<case>/case.toml), the regression overrides and
         boundary_order = 1 where the case does not set it (regression.py:273-276), as JSON.
The reference is also run here from U0 and must land on U_ref within the reference's tolerance 1e-8
(regression.py:305); the script stops if it does not.

    python tests/golden/make_golden.py [name ...]
"""
from __future__ import annotations

import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle_lib import Reference  # noqa: E402

REF_FIXTURES = "/root/reference/cpp/tests/reference"
TWO_PI = "6.283185307179586"

# case-file values of the keys of the path (cases/*/case.toml)
CASES = {
    "Taylor-Green Vortex 3D": dict(Lx=TWO_PI, Ly=TWO_PI, Lz=TWO_PI, nx=32, ny=32, nz=32, use_viscous="true", mu="6.25e-4",
                                   bc_xmin="periodic", bc_xmax="periodic", bc_ymin="periodic", bc_ymax="periodic",
                                   bc_zmin="periodic", bc_zmax="periodic", riemann_solver="HLLC (Low-Mach Thornber)",
                                   reconstruction="muscl", muscl_limiter="Van Leer", time_scheme="SSP Runge-Kutta 2", cfl=0.5),
    "Isentropic Vortex": dict(Lx=10.0, Ly=10.0, nx=80, ny=80, gamma=1.4, use_viscous="false", bc_xmin="periodic",
                              bc_xmax="periodic", bc_ymin="periodic", bc_ymax="periodic", riemann_solver="HLLC",
                              reconstruction="muscl", muscl_limiter="MC", time_scheme="SSP Runge-Kutta 2", cfl=0.5),
    "Sod Shock Tube": dict(Lx=1.0, Ly=0.1, nx=200, ny=20, use_viscous="false", bc_xmin="transmissive", bc_xmax="transmissive",
                           bc_ymin="wall", bc_ymax="wall", riemann_solver="HLLC", cfl=0.5),
    "Taylor-Green Vortex": dict(Lx=TWO_PI, Ly=TWO_PI, nx=64, ny=64, use_viscous="true", mu=0.02, bc_xmin="periodic",
                                bc_xmax="periodic", bc_ymin="periodic", bc_ymax="periodic",
                                riemann_solver="HLLC (Low-Mach Thornber)", reconstruction="muscl", muscl_limiter="Van Leer",
                                time_scheme="SSP Runge-Kutta 2", cfl=0.5),
    "Kelvin-Helmholtz": dict(Lx=14.1319, Ly=16.0, nx=96, ny=109, gamma=1.4, use_viscous="false", bc_xmin="periodic",
                             bc_xmax="periodic", bc_ymin="wall", bc_ymax="wall", riemann_solver="HLLC", reconstruction="muscl",
                             muscl_limiter="MC", time_scheme="SSP Runge-Kutta 2", cfl=0.4),
    "Shu-Osher Test": dict(Lx=10.0, Ly=0.1, nx=400, ny=2, gamma=1.4, use_viscous="false", bc_xmin="transmissive",
                           bc_xmax="transmissive", bc_ymin="wall", bc_ymax="wall", riemann_solver="HLLC", reconstruction="muscl",
                           muscl_limiter="MC", time_scheme="SSP Runge-Kutta 2", cfl=0.4),
    "Woodward-Colella Blast Waves": dict(Lx=1.0, Ly=0.01, nx=400, ny=4, gamma=1.4, use_viscous="false", bc_xmin="wall",
                                         bc_xmax="wall", bc_ymin="wall", bc_ymax="wall", riemann_solver="HLLC",
                                         reconstruction="muscl", muscl_limiter="MC", time_scheme="SSP Runge-Kutta 2", cfl=0.4),
    "Noh Implosion": dict(Lx=1.0, Ly=1.0, nx=200, ny=200, gamma="1.6666666666666667", use_viscous="false", bc_xmin="wall",
                          bc_ymin="wall", bc_xmax="transmissive", bc_ymax="transmissive", riemann_solver="HLLC",
                          reconstruction="muscl", muscl_limiter="MC", time_scheme="SSP Runge-Kutta 2", cfl=0.4),
    "Lid Driven Cavity": dict(Lx=1.0, Ly=1.0, nx=30, ny=30, use_viscous="true", mu="2.5e-3", bc_xmin="no_slip",
                              bc_xmax="no_slip", bc_ymin="no_slip", bc_ymax="moving_wall", u_in=1.0, rho_in=1.0, p_out=100.0,
                              riemann_solver="HLLC (Low-Mach Thornber)", time_scheme="SSP Runge-Kutta 3", cfl=0.4),
}

# regression.py:43-237, the cases on the uniform path
REGRESSION = {
    "tgv3d_weno5_rk3": ("Taylor-Green Vortex 3D", dict(nx=16, ny=16, nz=16, t_end=0.5, reconstruction="weno5",
                                                       time_scheme="SSP Runge-Kutta 3", cfl=0.4)),
    "tgv3d_weno3_rk3": ("Taylor-Green Vortex 3D", dict(nx=16, ny=16, nz=16, t_end=1.0, reconstruction="weno3",
                                                       time_scheme="SSP Runge-Kutta 3", cfl=0.4)),
    "tgv3d_16": ("Taylor-Green Vortex 3D", dict(nx=16, ny=16, nz=16, t_end=0.5, mu=0.01)),
    "hancock_tgv3d": ("Taylor-Green Vortex 3D", dict(nx=16, ny=24, nz=16, t_end=0.5, mu=0.01, reconstruction="muscl",
                                                     time_scheme="MUSCL-Hancock")),
    "vortex_weno5_rk3": ("Isentropic Vortex", dict(nx=48, ny=48, t_end=1.0, reconstruction="weno5",
                                                   time_scheme="SSP Runge-Kutta 3", cfl=0.4)),
    "vortex_weno3_rk3": ("Isentropic Vortex", dict(nx=48, ny=48, t_end=1.0, reconstruction="weno3",
                                                   time_scheme="SSP Runge-Kutta 3", cfl=0.4)),
    "isentropic_vortex": ("Isentropic Vortex", dict(nx=40, ny=40, t_end=1.0)),
    "sod_1st_euler": ("Sod Shock Tube", dict(nx=100, ny=4, t_end=0.05)),
    "sod_muscl_vl_rk2": ("Sod Shock Tube", dict(nx=100, ny=4, t_end=0.05, reconstruction="muscl", time_scheme="SSP Runge-Kutta 2")),
    "sod_muscl_mc_rk3_lm": ("Sod Shock Tube", dict(nx=80, ny=4, t_end=0.05, reconstruction="muscl", muscl_limiter="MC",
                                                   time_scheme="SSP Runge-Kutta 3", riemann_solver="HLLC (Low-Mach Thornber)")),
    "hancock_sod": ("Sod Shock Tube", dict(nx=100, ny=8, t_end=0.05, reconstruction="muscl", time_scheme="MUSCL-Hancock")),
    "taylor_green_visc": ("Taylor-Green Vortex", dict(nx=32, ny=32, t_end=0.5, mu=0.05)),
    "hancock_taylor_green": ("Taylor-Green Vortex", dict(nx=32, ny=32, t_end=0.5, mu=0.05, reconstruction="muscl",
                                                         time_scheme="MUSCL-Hancock")),
    "kelvin_helmholtz_32": ("Kelvin-Helmholtz", dict(nx=32, ny=36, t_end=4.0)),
    "sod3d_walls_z": ("Sod Shock Tube", dict(nx=40, ny=4, nz=4, Lz=0.1, t_end=0.05, bc_zmin="wall", bc_zmax="wall",
                                             reconstruction="muscl", time_scheme="SSP Runge-Kutta 2")),
    "sod3d_walls_z_b2": ("Sod Shock Tube", dict(nx=40, ny=4, nz=4, Lz=0.1, t_end=0.05, bc_zmin="wall", bc_zmax="wall",
                                                reconstruction="muscl", time_scheme="SSP Runge-Kutta 2", boundary_order=2)),
    "cavity_rk3_visc": ("Lid Driven Cavity", dict(nx=20, ny=20, t_end=0.2, reconstruction="muscl")),
    "cavity_weno5_b2": ("Lid Driven Cavity", dict(nx=20, ny=20, t_end=0.2, reconstruction="weno5",
                                                  time_scheme="SSP Runge-Kutta 3", boundary_order=2)),
    "cavity3d_noslip_z": ("Lid Driven Cavity", dict(nx=10, ny=10, nz=6, Lz=1.0, t_end=0.1, bc_zmin="no_slip",
                                                    bc_zmax="no_slip", reconstruction="muscl")),
    "cavity3d_noslip_z_b2": ("Lid Driven Cavity", dict(nx=10, ny=10, nz=6, Lz=1.0, t_end=0.1, bc_zmin="no_slip",
                                                       bc_zmax="no_slip", reconstruction="muscl", boundary_order=2)),
    "cavity3d_weno5_b2": ("Lid Driven Cavity", dict(nx=12, ny=12, nz=8, Lz=1.0, t_end=0.1, bc_zmin="no_slip",
                                                    bc_zmax="no_slip", reconstruction="weno5", time_scheme="SSP Runge-Kutta 3",
                                                    boundary_order=2)),
    # strong shocks, the EOS floors and periodic sides of a viscous cavity (regression.py:50, 52, 189-195, 212-213)
    "shu_osher": ("Shu-Osher Test", dict(nx=200, t_end=0.2)),
    "cavity_periodic_x": ("Lid Driven Cavity", dict(nx=16, ny=12, t_end=0.1, bc_xmin="Periodic", bc_xmax="Periodic")),
    "blast_waves_100": ("Woodward-Colella Blast Waves", dict(nx=100, ny=4, t_end=0.02)),
    "noh_40": ("Noh Implosion", dict(nx=40, ny=40, t_end=0.2)),
    "riemann2d_c3_32": ("riemann2d:3", dict(nx=32, ny=32, reconstruction="muscl", muscl_limiter="MC", time_scheme="SSP Runge-Kutta 2")),
    "riemann2d_c6_32": ("riemann2d:6", dict(nx=32, ny=32, reconstruction="muscl", muscl_limiter="MC", time_scheme="SSP Runge-Kutta 2")),
    "cavity_weno3_b2": ("Lid Driven Cavity", dict(nx=20, ny=20, t_end=0.2, reconstruction="weno3", time_scheme="SSP Runge-Kutta 3",
                                                  boundary_order=2)),
    # non-reflecting outlets (regression.py:174-182)
    "vortex_nrbc": ("Isentropic Vortex", dict(nx=40, ny=48, t_end=1.0, reconstruction="muscl", bc_xmin="transmissive",
                                              bc_xmax="nonreflecting")),
    "vortex_nrbc_xdecomp": ("Isentropic Vortex", dict(nx=64, ny=6, t_end=0.5, bc_xmin="nonreflecting", bc_xmax="nonreflecting",
                                                      nrbc_sigma=1.0)),
    "tgv3d_nrbc": ("Taylor-Green Vortex 3D", dict(nx=16, ny=16, nz=8, t_end=0.3, bc_xmin="transmissive", bc_xmax="nonreflecting",
                                                  bc_zmin="nonreflecting", bc_zmax="nonreflecting")),
}


# The four-quadrant Riemann problems of Lax and Liu (1998), configurations 3 and 6, as the reference's study writes
# them into a case file (cases/validation/verification/riemann_2d.py): t_end and the (rho, u, v, p) states of the
# north-east, north-west, south-west and south-east quadrants of [0, 1]^2, split at 0.5.
RIEMANN2D = {
    3: (0.3, (1.5, 0.0, 0.0, 1.5), (0.5323, 1.206, 0.0, 0.3), (0.138, 1.206, 1.206, 0.029), (0.5323, 0.0, 1.206, 0.3)),
    6: (0.3, (1.0, 0.75, -0.5, 1.0), (2.0, 0.75, 0.5, 1.0), (1.0, -0.75, 0.5, 1.0), (3.0, -0.75, -0.5, 1.0)),
}
for _k in RIEMANN2D:
    CASES[f"riemann2d:{_k}"] = dict(Lx=1.0, Ly=1.0, nx=100, ny=100, gamma=1.4, use_viscous="false", bc_xmin="transmissive",
                                    bc_xmax="transmissive", bc_ymin="transmissive", bc_ymax="transmissive", riemann_solver="HLLC",
                                    cfl=0.4, t_end=RIEMANN2D[_k][0])


def riemann2d_case(k, folder):
    """Configuration k as a case file in `folder` (for the reference's case loader); returns the folder."""
    t_end, ne, nw, sw, se = RIEMANN2D[k]

    def st(q):
        return f"{{ rho = {q[0]!r}, u = {q[1]!r}, v = {q[2]!r}, p = {q[3]!r} }}"

    def box(lo, hi, q):
        return f'[[initial.region]]\nshape = "box"\nmin = [{lo[0]!r}, {lo[1]!r}]\nmax = [{hi[0]!r}, {hi[1]!r}]\nstate = {st(q)}\n'

    text = (f'[case]\nname = "Riemann 2D config {k}"\n\n[mesh]\nLx = 1.0\nLy = 1.0\nnx = 100\nny = 100\n\n[boundaries]\n'
            'bc_xmin = "transmissive"\nbc_xmax = "transmissive"\nbc_ymin = "transmissive"\nbc_ymax = "transmissive"\n\n'
            f'[numerics]\nriemann_solver = "HLLC"\ncfl = 0.4\n\n[run]\nt_end = {t_end!r}\nsave_interval = 100000000\n\n'
            f'[initial]\nstate = {st(ne)}\n\n{box((0.0, 0.5), (0.5, 1.0), nw)}\n{box((0.0, 0.0), (0.5, 0.5), sw)}\n'
            f'{box((0.5, 0.0), (1.0, 0.5), se)}')
    os.makedirs(folder, exist_ok=True)
    with open(os.path.join(folder, "case.toml"), "w", newline="\n") as f:
        f.write(text)
    return folder


def keys_for(name):
    case, ov = REGRESSION[name]
    keys = dict(CASES[case])
    keys.update(ov)
    if "boundary_order" not in ov:
        keys["boundary_order"] = 1
    return case, {k: str(v) for k, v in keys.items()}


def main(names=None):
    ref = Reference()
    for name in names or REGRESSION:
        case, keys = keys_for(name)
        loader = case
        if case.startswith("riemann2d:"):
            loader = riemann2d_case(int(case.split(":")[1]), os.path.join(tempfile.mkdtemp(prefix="fvg_golden_"), case.replace(":", "_")))
        run = ref.create(loader, **keys)
        U0 = run.initial_field()
        stored = np.load(os.path.join(REF_FIXTURES, name + ".npz"))
        U_ref, step = stored["U"], int(stored["step"])
        run.set_state(U0)
        t, steps = run.run(float(keys["t_end"]))
        U = run.get_state()[:, : U_ref.shape[1]]
        scale = np.maximum(np.max(np.abs(U_ref), axis=0), 1e-30)
        diff = float(np.max(np.abs(U - U_ref) / scale))
        print(f"{name:24s} steps {steps}/{step} max rel diff {diff:.2e}")
        if steps != step or not diff < 1e-8:
            raise SystemExit(f"{name}: the reference run does not reproduce its stored state")
        np.savez_compressed(os.path.join(HERE, name + ".npz"), U0=U0, U_ref=U_ref, step=step,
                            keys=np.array(json.dumps(keys)), case=np.array(case))


if __name__ == "__main__":
    main(sys.argv[1:])
