"""Synthetic workloads of the benchmark and the tests.

* Case scripts: .hydroconf text on top of the defaults of
  examples/general.hydroconf (which hg_config_default holds) for the
  lid-driven cavity, the broken dam, the inlet/outlet channel and the
  Rayleigh-Taylor instability, plus option blocks (dt_auto, heat).

* Rayleigh-Taylor initial state: heavy fluid (phase 1) above
  y = 1 + 0.01 cos(2 pi x) + per-column noise U(-1e-4, 1e-4), with the
  cell-averaged volume fraction from the area above the interface
  (16-point Gauss-Legendre in x, exact clipping in y).
* Variable-coefficient Poisson operators: seeded two-phase volume-fraction
  snapshots at density ratio 1000, face coefficients k_f = A / (d rho_f),
  and a Neumann-compatible right-hand side (the discrete divergence of a
  smooth random velocity).
"""
import numpy as np

from . import capi


def cavity(n=64, dt=1e-3, viscosity=0.01, solver_p="gauss_seidel", solver_v="lu", extra=""):
    """Lid-driven cavity, unit square, lid velocity 1 (Re = 1/viscosity)."""
    return f"""
set int Nx {n}
set int Ny {n}
set double dt {dt}
set double viscosity_0 {viscosity}
set string condition_top "wall 1 0"
set string linear_solver_pressure {solver_p}
set string linear_solver_velocity {solver_v}
{extra}
"""


def dam(nx=256, ny=128, dt=1e-4, solver_p="gauss_seidel", solver_v="lu", extra=""):
    """Broken dam, 2x1 domain, water column 0.5 x 0.5 in the left corner:
    density ratio 1000, viscosity ratio 100, gravity -9.81."""
    return f"""
set vect A (0, 0)
set vect B (2, 1)
set int Nx {nx}
set int Ny {ny}
set int num_phases 2
set double density_0 1
set double density_1 1000
set double viscosity_0 1e-5
set double viscosity_1 1e-3
set vect gravity (0, -9.81)
set vect A1 (0, 0)
set vect B1 (0.5, 0.5)
set double dt {dt}
set string linear_solver_pressure {solver_p}
set string linear_solver_velocity {solver_v}
{extra}
"""


# dt_auto from a moving start (at rest every flux is 0 and dtm = 1e10); at
# cfl 0.4 the dam case blows up in the reference itself within two steps
DT_AUTO = """
set bool dt_auto 1
set double cfl 0.02
set double cfl_advection 0.01
set vect initial_velocity (0.2, 0.1)
"""


# heat: the bottom wall held at temperature 1 (the box holds the y = 0
# faces' centres), every other wall adiabatic; lu for the temperature
HEAT_BOTTOM = """
set bool heat_enable 1
set vect heat_box_lb (-1, -1)
set vect heat_box_rt (2, 0)
set double heat_box_temperature 1
set double temperature_initial 0.25
set double conductivity_0 0.02
set double conductivity_1 0.002
"""
# heat: the left wall held at temperature 2, first order in time, SOR for
# the temperature
HEAT_LEFT = """
set bool heat_enable 1
set string linear_solver_heat gauss_seidel
set vect heat_box_lb (0, 0)
set vect heat_box_rt (0, 1)
set double heat_box_temperature 2
set double conductivity_0 0.05
set double conductivity_1 0.001
set bool time_second_order_heat 0
set double heat_relaxation_factor 0.9
"""


def channel(nx=32, ny=16, dt=2e-3, extra=""):
    """Channel [0,2]x[0,1]: inflow 1 through the left side, outflow on the
    right, walls at the bottom and the top; two phases, a heavy block in the
    lower left so that both cross the inlet's and the outlet's faces."""
    return f"""
set vect A (0, 0)
set vect B (2, 1)
set int Nx {nx}
set int Ny {ny}
set int num_phases 2
set double density_0 1
set double density_1 5
set double viscosity_0 0.01
set double viscosity_1 0.02
set vect gravity (0, -1)
set vect A1 (0, 0)
set vect B1 (0.4, 0.3)
set double dt {dt}
set string condition_left "inlet 1 0"
set string condition_right "outlet"
{extra}
"""


def rt(nx=1024, ny=2048, dt=2e-4, solver_p="cg", solver_v="lu_rb", precision="fp64", extra=""):
    """Rayleigh-Taylor, [0,1]x[0,2], Atwood 0.5 (heavy 3 on top of light 1),
    g = (0,-1).  The interface is set through the partial-density fields."""
    return f"""
set vect A (0, 0)
set vect B (1, 2)
set int Nx {nx}
set int Ny {ny}
set int num_phases 2
set double density_0 1
set double density_1 3
set double viscosity_0 1e-3
set double viscosity_1 1e-3
set vect gravity (0, -1)
set int initial_volume_fraction_smooth_times 0
set double dt {dt}
set string linear_solver_pressure {solver_p}
set string linear_solver_velocity {solver_v}
set string precision_gpu {precision}
{extra}
"""


def rt_volume_fraction(nx, ny, lx=1.0, ly=2.0, amp=0.01, noise=1e-4, seed=2026):
    rng = np.random.default_rng(seed)
    xn = np.arange(nx + 1) / nx * lx
    yn = np.arange(ny + 1) / ny * ly
    col_noise = rng.uniform(-noise, noise, nx)
    gx, gw = np.polynomial.legendre.leggauss(16)
    x0, x1 = xn[:-1], xn[1:]
    xq = 0.5 * (x1 - x0)[:, None] * gx[None, :] + 0.5 * (x0 + x1)[:, None]   # (nx, 16)
    h = 0.5 * ly + amp * np.cos(2 * np.pi * xq) + col_noise[:, None]        # (nx, 16)
    alpha = np.zeros((ny, nx))
    dy = yn[1:] - yn[:-1]
    lo, hi = h.min() - 1e-12, h.max() + 1e-12
    j_lo = max(int(np.searchsorted(yn, lo)) - 1, 0)
    j_hi = min(int(np.searchsorted(yn, hi)) + 1, ny)
    alpha[j_hi:, :] = 1.0
    for j in range(j_lo, j_hi):
        above = np.clip(yn[j + 1] - np.maximum(h, yn[j]), 0.0, dy[j]) / dy[j]   # (nx, 16)
        alpha[j, :] = 0.5 * (above * gw[None, :]).sum(axis=1)
    return alpha.reshape(-1)


def set_rt_state(hydro, rho_light=1.0, rho_heavy=3.0, **kw):
    """Puts the RT interface into a Hydro (or Oracle) case and refreshes the
    derived fields (initial flux, properties)."""
    cfg = hydro.cfg
    a = rt_volume_fraction(cfg.nx, cfg.ny, lx=cfg.b[0] - cfg.a[0], ly=cfg.b[1] - cfg.a[1], **kw)
    hydro.set(capi.HG_FIELD_PARTIAL_DENSITY_0 + 1, a * rho_heavy)
    hydro.set(capi.HG_FIELD_PARTIAL_DENSITY_0, (1.0 - a) * rho_light)
    hydro.refresh()
    return a


def _smooth_once(a):
    p = np.pad(a, 1, mode="edge")
    return (4 * a + p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:]) / 8.0


def poisson_problem(n=512, seed=0, ratio=1000.0):
    """(kx, ky, rhs) of a variable-coefficient pressure system on an n x n
    unit square: rho = 1 + (ratio-1) alpha from random discs and a layer."""
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    xc = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(xc, xc)
    alpha = (Y < rng.uniform(0.2, 0.35)).astype(float)
    for _ in range(int(rng.integers(8, 17))):
        cx, cy = rng.uniform(0.1, 0.9, 2)
        rx, ry = rng.uniform(0.03, 0.12, 2)
        alpha[((X - cx) / rx) ** 2 + ((Y - cy) / ry) ** 2 < 1.0] = 1.0
    alpha = _smooth_once(alpha)
    rho = 1.0 + (ratio - 1.0) * alpha
    kx = np.zeros((n, n + 1))
    ky = np.zeros((n + 1, n))
    kx[:, 1:-1] = h / (h * 0.5 * (rho[:, :-1] + rho[:, 1:]))
    ky[1:-1, :] = h / (h * 0.5 * (rho[:-1, :] + rho[1:, :]))
    # rhs: divergence of a smooth random velocity (32 Fourier modes)
    xf = np.arange(n + 1) * h
    u = np.zeros((n, n + 1))
    v = np.zeros((n + 1, n))
    for _ in range(32):
        kx_, ky_ = rng.integers(1, 9, 2)
        a, b = rng.normal(size=2)
        ph1, ph2 = rng.uniform(0, 2 * np.pi, 2)
        u += a * np.sin(np.pi * kx_ * xf[None, :]) * np.cos(np.pi * ky_ * xc[:, None] + ph1)
        v += b * np.sin(np.pi * ky_ * xf[:, None]) * np.cos(np.pi * kx_ * xc[None, :] + ph2)
    div = (u[:, 1:] - u[:, :-1]) * h + (v[1:, :] - v[:-1, :]) * h
    rhs = div.reshape(-1)
    rhs = rhs - rhs.mean()
    return kx.reshape(-1), ky.reshape(-1), rhs


def c5_cases(count=64, seed=2026, solver_p="fcg_mg", solver_v="lu_rb", precision="mixed",
             sides=tuple(range(64, 513, 32))):
    """The batch workload: `count` independent cases, side n drawn from
    `sides` (64, 96, ..., 512), either a lid-driven cavity on the unit square
    (n x n) or a broken dam on the 2 x 1 box (n x n/2), physical parameters
    jittered by +-10 %; all seeded.  Returns .hydroconf texts."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.choice(sides))
        j = lambda v: v * (1.0 + rng.uniform(-0.1, 0.1))  # noqa: E731
        solvers = (f"set string linear_solver_pressure {solver_p}\n"
                   f"set string linear_solver_velocity {solver_v}\n"
                   f"set string precision_gpu {precision}\n")
        if rng.uniform() < 0.5:
            out.append(f"set int Nx {n}\nset int Ny {n}\nset double dt 1e-3\n"
                       f"set double viscosity_0 {j(0.01):.6g}\n"
                       f"set string condition_top \"wall {j(1.0):.6g} 0\"\n" + solvers)
        else:
            out.append(f"set vect A (0, 0)\nset vect B (2, 1)\nset int Nx {n}\nset int Ny {n // 2}\n"
                       f"set int num_phases 2\nset double density_0 {j(1.0):.6g}\n"
                       f"set double density_1 {j(1000.0):.6g}\n"
                       f"set double viscosity_0 {j(1e-5):.6g}\nset double viscosity_1 {j(1e-3):.6g}\n"
                       f"set vect gravity (0, {j(-9.81):.6g})\nset vect A1 (0, 0)\n"
                       f"set vect B1 (0.5, 0.5)\nset double dt 1e-4\n" + solvers)
    return out

# ---- 3D (hydro_gpu3d.h): vectors with three components; the keys whose
# general.hydroconf defaults have two (initial_velocity, box_A, box_B) are
# set with three, as the reference's 3D module reads three
_W3 = "set vect initial_velocity (0, 0, 0)\nset vect box_A (0, 0, 0)\nset vect box_B (0, 0, 0)\n"


def cavity3d(nx=16, ny=16, nz=16, dt=1e-3, viscosity=0.01, solver_p="jacobi",
             solver_v="jacobi", extra=""):
    """Lid-driven cavity in the unit cube: the lid (top, y = 1) moves in x."""
    return (f"set int Nx {nx}\nset int Ny {ny}\nset int Nz {nz}\nset vect A (0, 0, 0)\n"
            f"set vect B (1, 1, 1)\nset double dt {dt}\nset double viscosity_0 {viscosity}\n"
            f"set string condition_top \"wall 1 0 0\"\n"
            f"set string linear_solver_pressure {solver_p}\nset string linear_solver_velocity {solver_v}\n"
            f"set double lu_relaxed_relaxation_factor 0.9\n" + _W3 + extra)


def dam3d(nx=32, ny=8, nz=16, dt=1e-4, solver_p="jacobi", solver_v="jacobi", extra=""):
    """Broken dam in a 2 x 0.5 x 1 box, gravity -9.81 in z, density ratio
    1000, the water block 0.5 wide and 0.5 high in the corner across the
    whole depth (the reference's broken_dam_3d layout, smaller)."""
    return (f"set int Nx {nx}\nset int Ny {ny}\nset int Nz {nz}\nset vect A (0, 0, 0)\n"
            f"set vect B (2, 0.5, 1)\nset int num_phases 2\nset double density_0 1\n"
            f"set double density_1 1000\nset double viscosity_0 1e-5\nset double viscosity_1 1e-3\n"
            f"set vect gravity (0, 0, -9.81)\nset vect A1 (0, 0, 0)\nset vect B1 (0.5, 0.5, 0.5)\n"
            f"set double dt {dt}\nset string linear_solver_pressure {solver_p}\n"
            f"set string linear_solver_velocity {solver_v}\n"
            f"set double lu_relaxed_relaxation_factor 0.9\n" + _W3 + extra)
