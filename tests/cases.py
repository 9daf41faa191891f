"""Case scripts shared by the tests (.hydroconf text on top of the
defaults of examples/general.hydroconf, which hg_config_default holds)."""


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
